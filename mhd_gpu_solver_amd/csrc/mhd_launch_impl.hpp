// Launcher definitions shared by the per-precision translation units.
#pragma once

#include "mhd_launch.hpp"

#include <algorithm>

namespace mhdhip {

namespace detail {

inline dim3 stage_grid(const Tiling& tl)
{
    // one wave per (strip, row segment), four per block in row-band order; flat grid,
    // remapped in the kernel
    const long long tiles = static_cast<long long>(tl.strips) * tl.segs;
    return dim3(static_cast<unsigned>((tiles + kWavesPerBlock - 1) / kWavesPerBlock), 1, 1);
}

inline unsigned flat_blocks(long long work, int block, int cap)
{
    const long long b = (work + block - 1) / block;
    return static_cast<unsigned>(std::max<long long>(1, std::min<long long>(b, cap)));
}

template <int STAGE, int LIM, int NI, int REMAP, typename R, int FAST>
hipError_t launch_one(const StageBuffers& b, const Geom& g, const Physics& ph, const Tiling& tl,
                      int parity, hipStream_t s)
{
    // the kernel numbers its tiles in 32 bits (a grid of 2^31 tiles would hold over 10^11 cells)
    if (static_cast<long long>(tl.strips) * tl.segs > (1ll << 31)) {
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL((k_stage<STAGE, LIM, NI, REMAP, R, FAST>), stage_grid(tl), dim3(kBlock), 0, s, b,
                       g, ph, tl, parity);
    return hipGetLastError();
}

template <int STAGE, int LIM, typename R, int FAST>
hipError_t launch_ni_remap(bool ni, bool remap, const StageBuffers& b, const Geom& g,
                           const Physics& ph, const Tiling& tl, int parity, hipStream_t s)
{
    if (ni) {
        return remap ? launch_one<STAGE, LIM, 1, 1, R, FAST>(b, g, ph, tl, parity, s)
                     : launch_one<STAGE, LIM, 1, 0, R, FAST>(b, g, ph, tl, parity, s);
    }
    return remap ? launch_one<STAGE, LIM, 0, 1, R, FAST>(b, g, ph, tl, parity, s)
                 : launch_one<STAGE, LIM, 0, 0, R, FAST>(b, g, ph, tl, parity, s);
}

}  // namespace detail

template <typename R>
hipError_t launch_ctrl_init(Ctrl* ctrl, hipStream_t s)
{
    hipLaunchKernelGGL((k_ctrl_init<R>), dim3(1), dim3(64), 0, s, ctrl);
    return hipGetLastError();
}

template <typename R, int FAST>
hipError_t launch_cfl(const R* U, const Geom& g, const Physics& ph, const uint8_t* mask, Ctrl* ctrl,
                      hipStream_t s)
{
    const long long ncell = static_cast<long long>(g.nx) * g.ny;
    hipLaunchKernelGGL((k_cfl<R, FAST>), dim3(detail::flat_blocks(ncell, 256, 4096)), dim3(256), 0, s, U,
                       g, ph, mask, ctrl);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_ghost_fill(R* U, const Geom& g, const Gate& gt, hipStream_t s)
{
    const long long total = 2LL * g.ng * g.nxt + 2LL * g.ng * g.ny;
    hipLaunchKernelGGL((k_ghost_fill<R>), dim3(detail::flat_blocks(total, 256, 2048)), dim3(256), 0,
                       s, U, g, gt);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_wall_fill(R* U, const Geom& g, const long long* plan, int count, double gamma,
                            const Gate& gt, hipStream_t s)
{
    if (count <= 0) {
        return hipSuccess;
    }
    hipLaunchKernelGGL((k_wall_fill<R>), dim3(detail::flat_blocks(count, 256, 1 << 20)), dim3(256),
                       0, s, U, g, plan, count, gamma, gt);
    return hipGetLastError();
}

template <typename R, int FAST>
hipError_t launch_stage(int stage, int limiter, bool nonideal, bool remap, const StageBuffers& b,
                        const Geom& g, const Physics& ph, const Tiling& tl, int parity,
                        hipStream_t s)
{
    if (stage == 1) {
        return limiter == kMinmod
                   ? detail::launch_ni_remap<1, kMinmod, R, FAST>(nonideal, remap, b, g, ph, tl,
                                                                  parity, s)
                   : detail::launch_ni_remap<1, kMC, R, FAST>(nonideal, remap, b, g, ph, tl, parity,
                                                              s);
    }
    return limiter == kMinmod
               ? detail::launch_ni_remap<2, kMinmod, R, FAST>(nonideal, remap, b, g, ph, tl,
                                                              parity, s)
               : detail::launch_ni_remap<2, kMC, R, FAST>(nonideal, remap, b, g, ph, tl, parity, s);
}

namespace detail {

template <int STAGE, int LIM, int NI, typename R, int FAST>
hipError_t launch_batch_one(bool remap, const BatchTables& tb, int parity, int blocks,
                            hipStream_t s)
{
    if (remap) {
        hipLaunchKernelGGL((k_batch_stage<STAGE, LIM, NI, 1, R, FAST>), dim3(blocks), dim3(kBlock), 0, s,
                           tb, parity);
    } else {
        hipLaunchKernelGGL((k_batch_stage<STAGE, LIM, NI, 0, R, FAST>), dim3(blocks), dim3(kBlock), 0, s,
                           tb, parity);
    }
    return hipGetLastError();
}

template <int STAGE, typename R, int FAST>
hipError_t launch_batch_lim(int limiter, bool ni, bool remap, const BatchTables& tb, int parity,
                            int blocks, hipStream_t s)
{
    if (limiter == kMinmod) {
        return ni ? launch_batch_one<STAGE, kMinmod, 1, R, FAST>(remap, tb, parity, blocks, s)
                  : launch_batch_one<STAGE, kMinmod, 0, R, FAST>(remap, tb, parity, blocks, s);
    }
    return ni ? launch_batch_one<STAGE, kMC, 1, R, FAST>(remap, tb, parity, blocks, s)
              : launch_batch_one<STAGE, kMC, 0, R, FAST>(remap, tb, parity, blocks, s);
}

}  // namespace detail

template <typename R, int FAST>
hipError_t launch_batch_stage(int stage, int limiter, bool nonideal, bool remap,
                              const BatchTables& tb, int parity, int blocks, hipStream_t s)
{
    return stage == 1
               ? detail::launch_batch_lim<1, R, FAST>(limiter, nonideal, remap, tb, parity, blocks,
                                                      s)
               : detail::launch_batch_lim<2, R, FAST>(limiter, nonideal, remap, tb, parity, blocks,
                                                      s);
}

template <typename R>
hipError_t launch_batch_fill(const BatchTables& tb, int parity, int stage, int slots,
                             hipStream_t s)
{
    hipLaunchKernelGGL((k_batch_fill<R>), dim3(slots), dim3(256), 0, s, tb, parity, stage);
    return hipGetLastError();
}

namespace detail {

template <int STAGE, int LIM, int NI, typename R, int FAST>
hipError_t batch_occupancy(bool remap, int* n)
{
    return remap ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                       n, k_batch_stage<STAGE, LIM, NI, 1, R, FAST>, kBlock, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                       n, k_batch_stage<STAGE, LIM, NI, 0, R, FAST>, kBlock, 0);
}

template <int STAGE, typename R, int FAST>
hipError_t batch_occupancy_lim(int limiter, bool nonideal, bool remap, int* n)
{
    if (limiter == kMinmod) {
        return nonideal ? batch_occupancy<STAGE, kMinmod, 1, R, FAST>(remap, n)
                        : batch_occupancy<STAGE, kMinmod, 0, R, FAST>(remap, n);
    }
    return nonideal ? batch_occupancy<STAGE, kMC, 1, R, FAST>(remap, n)
                    : batch_occupancy<STAGE, kMC, 0, R, FAST>(remap, n);
}

}  // namespace detail

template <typename R, int FAST>
int batch_blocks_per_cu(int stage, int limiter, bool nonideal, bool remap)
{
    int n = 0;
    const hipError_t e =
        stage == 1 ? detail::batch_occupancy_lim<1, R, FAST>(limiter, nonideal, remap, &n)
                   : detail::batch_occupancy_lim<2, R, FAST>(limiter, nonideal, remap, &n);
    return (e == hipSuccess && n > 0) ? n : 1;
}

template <typename R>
hipError_t launch_max_div_b(const R* U, const Geom& g, double inv_2dx, double inv_2dy,
                            unsigned long long* out, hipStream_t s)
{
    const hipError_t e = hipMemsetAsync(out, 0, sizeof(unsigned long long), s);
    if (e != hipSuccess) {
        return e;
    }
    const long long n = static_cast<long long>(g.nx) * g.ny;
    const int blocks = static_cast<int>(std::max<long long>(
        1, std::min<long long>(kTotalsBlocks, (n + kTotalsBlock - 1) / kTotalsBlock)));
    hipLaunchKernelGGL((k_max_div_b<R>), dim3(blocks), dim3(kTotalsBlock), 0, s, U, g, inv_2dx,
                       inv_2dy, out);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_totals(const R* U, const Geom& g, DD* partial, double* out, hipStream_t s)
{
    hipLaunchKernelGGL((k_totals_partial<R>), dim3(kTotalsBlocks), dim3(kTotalsBlock), 0, s, U, g,
                       partial);
    hipLaunchKernelGGL((k_totals_final<R>), dim3(1), dim3(kNC * kWave), 0, s, partial,
                       kTotalsBlocks, out);
    return hipGetLastError();
}

template <typename RA, typename RB>
hipError_t launch_rel_l1(const RA* A, const RB* B, const Geom& g, DD* partial, double* out,
                         hipStream_t s)
{
    const long long n = static_cast<long long>(g.nx) * g.ny;
    const int blocks = static_cast<int>(std::max<long long>(
        1, std::min<long long>(kTotalsBlocks, (n + kTotalsBlock - 1) / kTotalsBlock)));
    hipLaunchKernelGGL((k_l1_partial<RA, RB>), dim3(blocks), dim3(kTotalsBlock), 0, s, A, B, g,
                       partial);
    hipLaunchKernelGGL(k_l1_final<kNC>, dim3(1), dim3(kNC * kWave), 0, s, partial, blocks,
                       out);
    return hipGetLastError();
}

template <typename R, int FAST>
int stage_blocks_per_cu(int limiter, bool nonideal, bool remap)
{
    int n = 0;
    hipError_t e;
    auto q = [&](auto kern) { return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kern, kBlock, 0); };
    if (limiter == kMinmod) {
        e = nonideal ? (remap ? q(k_stage<2, kMinmod, 1, 1, R, FAST>)
                              : q(k_stage<2, kMinmod, 1, 0, R, FAST>))
                     : (remap ? q(k_stage<2, kMinmod, 0, 1, R, FAST>)
                              : q(k_stage<2, kMinmod, 0, 0, R, FAST>));
    } else {
        e = nonideal ? (remap ? q(k_stage<2, kMC, 1, 1, R, FAST>) : q(k_stage<2, kMC, 1, 0, R, FAST>))
                     : (remap ? q(k_stage<2, kMC, 0, 1, R, FAST>) : q(k_stage<2, kMC, 0, 0, R, FAST>));
    }
    return (e == hipSuccess && n > 0) ? n : 1;
}

// The single-run launchers of one precision, and the batch engine's (separate, so that a
// precision's two sets can be compiled with different contraction flags).
#define MHDHIP_INSTANTIATE(R) MHDHIP_INSTANTIATE_SINGLE(R) MHDHIP_INSTANTIATE_BATCH(R)
#define MHDHIP_INSTANTIATE_SINGLE(R)                                                             \
    template hipError_t launch_ctrl_init<R>(Ctrl*, hipStream_t);                                 \
    template hipError_t launch_cfl<R>(const R*, const Geom&, const Physics&, const uint8_t*,     \
                                      Ctrl*, hipStream_t);                                       \
    template hipError_t launch_ghost_fill<R>(R*, const Geom&, const Gate&, hipStream_t);         \
    template hipError_t launch_wall_fill<R>(R*, const Geom&, const long long*, int, double,      \
                                            const Gate&, hipStream_t);                           \
    template hipError_t launch_stage<R>(int, int, bool, bool, const StageBuffers&, const Geom&,  \
                                        const Physics&, const Tiling&, int, hipStream_t);         \
    template int stage_blocks_per_cu<R>(int, bool, bool);                                        \
    template hipError_t launch_totals<R>(const R*, const Geom&, DD*, double*, hipStream_t);     \
    template hipError_t launch_max_div_b<R>(const R*, const Geom&, double, double,              \
                                            unsigned long long*, hipStream_t);
#define MHDHIP_INSTANTIATE_BATCH(R)                                                              \
    template hipError_t launch_batch_stage<R>(int, int, bool, bool, const BatchTables&, int, int, \
                                              hipStream_t);                                       \
    template hipError_t launch_batch_fill<R>(const BatchTables&, int, int, int, hipStream_t);      \
    template int batch_blocks_per_cu<R>(int, int, bool, bool);

// The stepping launchers of the fast fp64 build (mhd_fp64fast.hip).
#define MHDHIP_INSTANTIATE_FAST(R)                                                               \
    template hipError_t launch_cfl<R, 1>(const R*, const Geom&, const Physics&, const uint8_t*,  \
                                         Ctrl*, hipStream_t);                                    \
    template hipError_t launch_stage<R, 1>(int, int, bool, bool, const StageBuffers&,            \
                                           const Geom&, const Physics&, const Tiling&, int,      \
                                           hipStream_t);                                         \
    template hipError_t launch_batch_stage<R, 1>(int, int, bool, bool, const BatchTables&, int,  \
                                                 int, hipStream_t);                              \
    template int batch_blocks_per_cu<R, 1>(int, int, bool, bool);                                \
    template int stage_blocks_per_cu<R, 1>(int, bool, bool);

}  // namespace mhdhip
