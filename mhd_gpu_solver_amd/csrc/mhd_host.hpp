// Host-only helpers of the hip backend (no device needed).
#pragma once

#include "mhd_hip.h"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace mhdhip {

// Writes the conserved interior of a named initial condition; throws std::invalid_argument for
// an unknown name.
int init_ic(const char* name, const mhd_hip_grid& g, double gamma, double b0, uint64_t seed,
            double* const comps[9]);

// Solid-cell fill plan as (dst, src or -1, axis) triples; returns the number of solid cells.
size_t wall_plan(const mhd_hip_grid& g, const uint8_t* mask, std::vector<long long>& plan);

}  // namespace mhdhip
