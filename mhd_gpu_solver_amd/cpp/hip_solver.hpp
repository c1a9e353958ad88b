// The hip backend's entry points on the reference's side (include/mhd/gpu/device_solver.hpp:
// the declarations a maintainer adds next to cuda_available / make_cuda_solver, :104-135).
#pragma once

#include "mhd/gpu/device_solver.hpp"

#include <memory>

namespace mhd {

bool hip_available();
std::unique_ptr<DeviceSolver> make_hip_solver(const Grid2D& grid, Precision precision);

}  // namespace mhd
