// Host-side launchers of the step kernels, one explicit instantiation per precision:
// mhd_fp64.hip (double, compiled with -ffp-contract=off) and mhd_fp32.hip (float, contraction on).
#pragma once

#include "mhd_batch_kernels.hpp"

#include <hip/hip_runtime.h>

namespace mhdhip {

// Wave/segment tiling of the fused stage kernel for an nx x ny interior.
Tiling make_tiling(int nx, int ny, int seg_rows);
int default_seg_rows(int nx, int ny);

template <typename R>
hipError_t launch_ctrl_init(Ctrl* ctrl, hipStream_t s);

// FAST = 1: the stepping kernels of the fast fp64 build (R = double), defined in
// mhd_fp64fast.hip; FAST = 0 everywhere else.
template <typename R, int FAST = 0>
hipError_t launch_cfl(const R* U, const Geom& g, const Physics& ph, const uint8_t* mask, Ctrl* ctrl,
                      hipStream_t s);

template <typename R>
hipError_t launch_ghost_fill(R* U, const Geom& g, const Gate& gt, hipStream_t s);

template <typename R>
hipError_t launch_wall_fill(R* U, const Geom& g, const long long* plan, int count, double gamma,
                            const Gate& gt, hipStream_t s);

template <typename R, int FAST = 0>
hipError_t launch_stage(int stage, int limiter, bool nonideal, bool remap, const StageBuffers& b,
                        const Geom& g, const Physics& ph, const Tiling& tl, int parity,
                        hipStream_t s);

// Compensated totals of the nine conserved fields over the interior: `partial` must hold
// kTotalsBlocks * 9 DD values, `out` 9 doubles.
constexpr int kTotalsBlocks = 512;
template <typename R>
hipError_t launch_totals(const R* U, const Geom& g, DD* partial, double* out, hipStream_t s);

// Relative L1 distance per component of A from the reference B (same grid) over the interior:
// `partial` must hold kTotalsBlocks * 18 DD values, `out` 9 doubles. Instantiated for every
// pairing of float and double.
template <typename RA, typename RB>
hipError_t launch_rel_l1(const RA* A, const RB* B, const Geom& g, DD* partial, double* out,
                         hipStream_t s);

// The same on type-erased states: a64 / b64 tell whether A / B store fp64 (else fp32).
inline hipError_t launch_rel_l1_any(const void* A, bool a64, const void* B, bool b64,
                                    const Geom& g, DD* partial, double* out, hipStream_t s)
{
    if (a64 && b64) {
        return launch_rel_l1(static_cast<const double*>(A), static_cast<const double*>(B), g,
                             partial, out, s);
    }
    if (a64) {
        return launch_rel_l1(static_cast<const double*>(A), static_cast<const float*>(B), g,
                             partial, out, s);
    }
    if (b64) {
        return launch_rel_l1(static_cast<const float*>(A), static_cast<const double*>(B), g,
                             partial, out, s);
    }
    return launch_rel_l1(static_cast<const float*>(A), static_cast<const float*>(B), g, partial,
                         out, s);
}

// max |div B| of the interior by central differences (stored ghosts used) into *out.
template <typename R>
hipError_t launch_max_div_b(const R* U, const Geom& g, double inv_2dx, double inv_2dy,
                            unsigned long long* out, hipStream_t s);

// Workgroups of k_stage (stage 2, the larger) that fit on one CU at once (occupancy query).
template <typename R, int FAST = 0>
int stage_blocks_per_cu(int limiter, bool nonideal, bool remap);

// Batched engine: one persistent launch of `blocks` workgroups per stage and kernel variant
// (remap: ghosts addressed at their sources; otherwise read from memory after k_batch_fill).
template <typename R, int FAST = 0>
hipError_t launch_batch_stage(int stage, int limiter, bool nonideal, bool remap,
                              const BatchTables& tb, int parity, int blocks, hipStream_t s);

// Ghost and solid-wall fill of the listed instances' stage input, one workgroup per slot of
// `slots` (the engine's slot count; workgroups past the list's length return at once).
template <typename R>
hipError_t launch_batch_fill(const BatchTables& tb, int parity, int stage, int slots,
                             hipStream_t s);

// Workgroups of k_batch_stage of this stage that fit on one CU at once (occupancy query).
template <typename R, int FAST = 0>
int batch_blocks_per_cu(int stage, int limiter, bool nonideal, bool remap);

}  // namespace mhdhip
