// HIP kernels of one SSP-RK2 step of the GLM-MHD solver (gfx950, wave64).
//
// Kernels
//   k_ctrl_init   zero the device control block (no memset nodes, graph-capture safe)
//   k_cfl         max signal speed over fluid interior cells -> ctrl.chmax[0]
//   k_ghost_fill  all ghost cells from their interior sources (x-then-y fill, corners exact)
//   k_wall_fill   reflective slip fill of solid cells from a precomputed neighbour plan
//   k_stage<S>    fused RK stage S in {1,2}: MUSCL slopes, HLL x/y faces, non-ideal E,
//                 flux difference, GLM psi damping, solid zero, RK combine; stage 2 also
//                 reduces the next step's CFL signal speed (wave shuffle + one atomic/wave).
//
// Step bookkeeping stays on the device: ctrl holds (t, steps, c_h, chmax) double-buffered by
// step parity; every stage kernel recomputes dt from chmax[p] with the reference formula
// (include/mhd/cfl.hpp:77-89, include/mhd/time_integrator.hpp:89) and returns at once when the
// step is inactive (t >= t_end or steps >= max_steps), so surplus steps in a launched batch
// are no-ops and a batch can run ahead of the host.
#pragma once

#include "mhd_batch_plan.hpp"
#include "mhd_physics.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace mhdhip {

constexpr int kWave = 64;
constexpr int kStripCells = kWave - 4;  // cells per wave strip: 2 halo lanes left, 2 right
constexpr int kHalfCells = kWave / 2 - 4;  // cells per half-wave strip (narrow tail strips)
constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWave * kWavesPerBlock;

// Index of the calling wave in its workgroup. Lanes 64w..64w+63 form wave w, so the value is
// the same on every lane; reading it through readfirstlane tells the compiler so, and what is
// derived from it (strip, rows, row addresses, loop bounds) stays in scalar registers.
__device__ __forceinline__ int wave_in_block()
{
    return __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
}

// Neighbour exchange across the wave through the whole-wave DPP shifts of GFX9 (wave_shr:1,
// wave_shl:1): one VALU move, often folded into its consumer, instead of an LDS permute.
// from_left(x) on lane i is x of lane i - 1, from_right(x) x of lane i + 1; the edge lane with
// no source (lane 0, resp. 63) reads 0. Callers keep all 64 lanes active, and nothing a cell
// lane stores depends on an edge lane's zero (see kXFaceLanes).
template <int CTRL>
__device__ __forceinline__ int dpp_move(int x)
{
    return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ float dpp_move(float x)
{
    return __builtin_bit_cast(float, dpp_move<CTRL>(__builtin_bit_cast(int, x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x)
{
    const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = static_cast<unsigned>(dpp_move<CTRL>(static_cast<int>(v & 0xffffffffu)));
    const unsigned hi = static_cast<unsigned>(dpp_move<CTRL>(static_cast<int>(v >> 32)));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}
// x-faces whose fluxes cell lanes 2..61 use: the left faces of lanes 2..62. Faces of lanes 0,
// 1 and 63 see zero-filled neighbours and are never used; wave-uniform branch decisions on
// x-faces are taken over these lanes only.
constexpr unsigned long long kXFaceLanes = ((1ull << 61) - 1ull) << 2;
// the same for a wave split into two half-wave strips: lanes 2..30 and 34..62
constexpr unsigned long long kXFaceLanesHalf =
    (((1ull << 29) - 1ull) << 2) | (((1ull << 29) - 1ull) << 34);

constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;
template <typename R>
__device__ __forceinline__ R from_left(R x)
{
    return dpp_move<kDppWaveShr1>(x);
}
template <typename R>
__device__ __forceinline__ R from_right(R x)
{
    return dpp_move<kDppWaveShl1>(x);
}

struct Ctrl {
    double t[2];
    double ch_last[2];
    unsigned long long chmax[2];
    int steps[2];
    int bad_dt;  // set when an active step got dt <= 0 (the reference throws there)
    int pad;
};

struct Geom {
    int nx, ny, ng, nxt, nyt;
    long long plane;  // elements per component plane, nxt * nyt
    int bc_x, bc_y;   // 0 periodic, 2 outflow
};

// A 64-bit value from its two 32-bit words. Both words are unsigned, so that the low word
// cannot sign-extend into the high one.
__host__ __device__ constexpr unsigned long long join_words(unsigned hi, unsigned lo)
{
    return (static_cast<unsigned long long>(hi) << 32) | static_cast<unsigned long long>(lo);
}
static_assert(join_words(0x1u, 0x80000000u) == 0x180000000ull &&
                  join_words(0x7fffu, 0xfffffff0u) == 0x7ffffffffff0ull,
              "a low word with bit 31 set stays in the low half");

// A wave-uniform value read into scalar registers (readfirstlane returns an int per word).
__device__ __forceinline__ int uniform_i32(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long x)
{
    return join_words(
        static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(x >> 32))),
        static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(x & 0xffffffffu))));
}

// A copy of g read into scalar registers (g is wave-uniform but may have been read by vector
// loads).
__device__ __forceinline__ Geom uniform_geom(const Geom& g)
{
    Geom u;
    u.nx = uniform_i32(g.nx);
    u.ny = uniform_i32(g.ny);
    u.ng = uniform_i32(g.ng);
    u.nxt = uniform_i32(g.nxt);
    u.nyt = uniform_i32(g.nyt);
    u.plane = static_cast<long long>(uniform_u64(static_cast<unsigned long long>(g.plane)));
    u.bc_x = uniform_i32(g.bc_x);
    u.bc_y = uniform_i32(g.bc_y);
    return u;
}

struct Physics {
    double gamma, gm1, inv_gm1;
    double cfl, inv_dx, inv_dy, inv_2dx, inv_2dy;
    double glm_alpha, h;
    double dt_diff;  // cfl h^2 / (4 eta_max); only used when has_diff
    int has_diff;
    double t_end;
    int max_steps;
    NonIdeal ni;
};

struct Tiling {
    int strips;    // 60-column strips across nx
    int seg_rows;  // rows marched by one wave
    int segs;      // row segments across ny
};

struct StepState {
    double t, dt, c_h;
    int steps;
    bool active;
};

// dt for the step with parity p: include/mhd/cfl.hpp:77-89, time_integrator.hpp:85-90
__device__ __forceinline__ StepState step_state(const Ctrl* ctrl, const Physics& ph, int p)
{
    StepState s;
    s.t = ctrl->t[p];
    s.steps = ctrl->steps[p];
    s.c_h = __longlong_as_double(static_cast<long long>(ctrl->chmax[p]));
    double dt = (s.c_h <= 0.0) ? __builtin_huge_val() : ph.cfl / (s.c_h * ph.inv_dx + s.c_h * ph.inv_dy);
    if (ph.has_diff) {
        dt = vmin(dt, ph.dt_diff);
    }
    s.dt = vmin(dt, ph.t_end - s.t);
    // include/mhd/time_integrator.hpp:42-44: a step with dt <= 0 is an error, never executed
    s.active = (s.t < ph.t_end) && (s.steps < ph.max_steps) && (s.dt > 0.0);
    return s;
}

// step not yet finished by time or count (an inactive step with this true had dt <= 0)
__device__ __forceinline__ bool ctrl_pending(const Ctrl* ctrl, const Physics& ph, int p)
{
    return ctrl->t[p] < ph.t_end && ctrl->steps[p] < ph.max_steps;
}

template <typename R>
__device__ __forceinline__ R wave_max(R v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const R o = __shfl_xor(v, off);
        v = vmax(v, o);
    }
    return v;
}

__device__ __forceinline__ void atomic_max_speed(unsigned long long* slot, double v)
{
    // non-negative doubles order like their bit patterns
    atomicMax(slot, static_cast<unsigned long long>(__double_as_longlong(v)));
}

template <typename R>
__global__ void k_ctrl_init(Ctrl* ctrl)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl->t[0] = 0.0;
        ctrl->t[1] = 0.0;
        ctrl->ch_last[0] = 0.0;
        ctrl->ch_last[1] = 0.0;
        ctrl->chmax[0] = 0ull;
        ctrl->chmax[1] = 0ull;
        ctrl->steps[0] = 0;
        ctrl->steps[1] = 0;
        ctrl->bad_dt = 0;
        ctrl->pad = 0;
    }
}

// ---- CFL over the fluid interior: include/mhd/cfl.hpp:46-71 --------------------------------
// FAST (every stepping kernel): 1 in the fast fp64 translation unit only, so that its kernels,
// built with other arithmetic, never share a symbol with the exact ones.
template <typename R, int FAST = 0>
__global__ __launch_bounds__(256) void k_cfl(const R* __restrict__ U, Geom g, Physics ph,
                                             const uint8_t* __restrict__ mask, Ctrl* ctrl)
{
    static_assert(FAST == MHD_FAST_FP64, "stepping kernel built in the wrong translation unit");
    const long long ncell = static_cast<long long>(g.nx) * g.ny;
    R local = R(0);
    const R gamma = static_cast<R>(ph.gamma);
    const R gm1 = static_cast<R>(ph.gm1);
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < ncell;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        if (mask != nullptr && mask[k] == 0) {
            continue;
        }
        const int jj = static_cast<int>(k / g.nx);
        const int ii = static_cast<int>(k - static_cast<long long>(jj) * g.nx);
        const long long cell = static_cast<long long>(jj + g.ng) * g.nxt + (ii + g.ng);
        R u[kNC];
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            u[c] = U[c * g.plane + cell];
        }
        local = vmax(local, cell_max_speed(u, gamma, gm1));
    }
    local = wave_max(local);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomic_max_speed(&ctrl->chmax[0], static_cast<double>(local));
    }
}

// ---- ghost fill: include/boundary_conditions/boundary_conditions.hpp:33-103 ---------------
// A ghost's value after the x pass then the y pass is the interior cell at the mapped (i, j).
// Gate: the ghost and wall fills run inside a step only while that step is active (so surplus
// no-op steps leave the state untouched); parity < 0 runs unconditionally.
struct Gate {
    const Ctrl* ctrl;
    double t_end;
    int max_steps;
    int parity;
};

__device__ __forceinline__ bool gate_open(const Gate& gt)
{
    if (gt.parity < 0) {
        return true;
    }
    return gt.ctrl->t[gt.parity] < gt.t_end && gt.ctrl->steps[gt.parity] < gt.max_steps;
}

// Ghost cell t of the 2 ng (nxt + ny) ghost cells of a grid, in the order bottom band, top
// band, then the left and right columns of the interior rows, takes its interior source.
template <typename R>
__device__ __forceinline__ void ghost_fill_cell(R* __restrict__ U, const Geom& g, long long t)
{
    const long long band = static_cast<long long>(g.ng) * g.nxt;
    int i, j;
    if (t < band) {
        j = static_cast<int>(t / g.nxt);
        i = static_cast<int>(t - static_cast<long long>(j) * g.nxt);
    } else if (t < 2 * band) {
        const long long t2 = t - band;
        j = g.ng + g.ny + static_cast<int>(t2 / g.nxt);
        i = static_cast<int>(t2 % g.nxt);
    } else {
        const long long t3 = t - 2 * band;
        const int w = 2 * g.ng;
        j = g.ng + static_cast<int>(t3 / w);
        const int k = static_cast<int>(t3 % w);
        i = k < g.ng ? k : g.nx + k;
    }
    const int si = ghost_source(i, g.nx, g.ng, g.bc_x);
    const int sj = ghost_source(j, g.ny, g.ng, g.bc_y);
    const long long dst = static_cast<long long>(j) * g.nxt + i;
    const long long src = static_cast<long long>(sj) * g.nxt + si;
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        U[c * g.plane + dst] = U[c * g.plane + src];
    }
}

__device__ __host__ __forceinline__ long long ghost_cells(const Geom& g)
{
    return 2 * static_cast<long long>(g.ng) * g.nxt + static_cast<long long>(g.ny) * 2 * g.ng;
}

template <typename R>
__global__ __launch_bounds__(256) void k_ghost_fill(R* __restrict__ U, Geom g, Gate gt)
{
    if (!gate_open(gt)) {
        return;
    }
    const long long total = ghost_cells(g);
    for (long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<long long>(gridDim.x) * blockDim.x) {
        ghost_fill_cell(U, g, t);
    }
}

// ---- solid wall fill: include/mhd/mask.hpp:85-149 ------------------------------------------
// plan[3k] = destination cell, plan[3k+1] = fluid source cell or -1, plan[3k+2] = axis (0 x, 1 y)
// Sources are always fluid cells, so every entry is independent of the others.
// Solid cell k of the plan takes its reflected fluid neighbour (or the quiescent default).
template <typename R>
__device__ __forceinline__ void wall_fill_cell(R* __restrict__ U, const Geom& g,
                                               const long long* __restrict__ plan, int k,
                                               double gamma)
{
    const long long dst = plan[3 * k];
    const long long src = plan[3 * k + 1];
    const int axis = static_cast<int>(plan[3 * k + 2]);
    double W[kNC];
    if (src >= 0) {
        double S[kNC];
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            S[c] = static_cast<double>(U[c * g.plane + src]);
        }
        // to_primitive: include/mhd/mhd_physics.hpp:23-39 (divides by rho)
        W[0] = S[0];
        W[1] = S[1] / S[0];
        W[2] = S[2] / S[0];
        W[3] = S[3] / S[0];
        W[5] = S[5];
        W[6] = S[6];
        W[7] = S[7];
        W[8] = S[8];
        const double e_kin = 0.5 * (S[1] * S[1] + S[2] * S[2] + S[3] * S[3]) / S[0];
        const double e_mag = 0.5 * (S[5] * S[5] + S[6] * S[6] + S[7] * S[7]);
        W[4] = (gamma - 1.0) * (S[4] - e_kin - e_mag);
        if (axis == 0) {
            W[1] = -W[1];
        } else {
            W[2] = -W[2];
        }
        if (W[0] < 1e-12) {
            W[0] = 1e-12;
        }
        if (W[4] < 1e-12) {
            W[4] = 1e-12;
        }
    } else {
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            W[c] = 0.0;
        }
        W[0] = 1.0;
        W[4] = 1.0;
    }
    // to_conservative: include/mhd/mhd_physics.hpp:41-59
    double O[kNC];
    O[0] = W[0];
    O[1] = W[0] * W[1];
    O[2] = W[0] * W[2];
    O[3] = W[0] * W[3];
    O[5] = W[5];
    O[6] = W[6];
    O[7] = W[7];
    O[8] = W[8];
    const double ek = 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
    const double em = 0.5 * (W[5] * W[5] + W[6] * W[6] + W[7] * W[7]);
    O[4] = W[4] / (gamma - 1.0) + ek + em;
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        U[c * g.plane + dst] = static_cast<R>(O[c]);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_wall_fill(R* __restrict__ U, Geom g,
                                                   const long long* __restrict__ plan, int count,
                                                   double gamma, Gate gt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count || !gate_open(gt)) {
        return;
    }
    wall_fill_cell(U, g, plan, k, gamma);
}

// ---- compensated totals of the conserved fields -----------------------------------------
// Sums of each component over the interior cells in double-double arithmetic (Knuth's
// TwoSum), so the totals used for mass/energy drift carry ~106 bits whatever the storage
// precision. Pass 1 writes one (hi, lo) pair per block and component; pass 2 reduces the
// pairs in block order. No atomics: the result is deterministic.
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, double b)
{
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b - bb);
    const double lo = a.lo + err;
    const double hi = s + lo;
    return DD{hi, lo - (hi - s)};
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    DD r = dd_add(a, b.hi);
    return dd_add(r, b.lo);
}

constexpr int kTotalsBlock = 256;

template <typename R>
__global__ __launch_bounds__(kTotalsBlock) void k_totals_partial(const R* __restrict__ U, Geom g,
                                                                 DD* __restrict__ partial)
{
    __shared__ DD red[kNC][kTotalsBlock / kWave];
    const long long ncell = static_cast<long long>(g.nx) * g.ny;
    DD acc[kNC];
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        acc[c] = DD{0.0, 0.0};
    }
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < ncell;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int jj = static_cast<int>(k / g.nx);
        const int ii = static_cast<int>(k - static_cast<long long>(jj) * g.nx);
        const long long cell = static_cast<long long>(jj + g.ng) * g.nxt + (ii + g.ng);
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            acc[c] = dd_add(acc[c], static_cast<double>(U[c * g.plane + cell]));
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        DD v = acc[c];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const DD o{__shfl_down(v.hi, off), __shfl_down(v.lo, off)};
            v = dd_add(v, o);
        }
        if (lane == 0) {
            red[c][wave] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < kNC) {
        DD v = red[threadIdx.x][0];
        for (int w = 1; w < kTotalsBlock / kWave; ++w) {
            v = dd_add(v, red[threadIdx.x][w]);
        }
        partial[blockIdx.x * kNC + threadIdx.x] = v;
    }
}

// max |dBx/dx + dBy/dy| over the interior cells by central differences of the stored state,
// ghosts included, evaluated in fp64 without contraction:
// include/initial_condition/inititial_condition.hpp:273-287. The running max follows
// std::max (a NaN never replaces it); the values are >= 0, so their bit patterns order like
// the values and one unsigned atomic max per wave combines them in any order. `out` starts at 0.
template <typename R>
__global__ __launch_bounds__(kTotalsBlock) void k_max_div_b(const R* __restrict__ U, Geom g,
                                                            double inv_2dx, double inv_2dy,
                                                            unsigned long long* out)
{
#pragma clang fp contract(off)
    const R* bx = U + BX * g.plane;
    const R* by = U + BY * g.plane;
    const long long n = static_cast<long long>(g.nx) * g.ny;
    double m = 0.0;
    for (long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int j = static_cast<int>(k / g.nx);
        const int i = static_cast<int>(k - static_cast<long long>(j) * g.nx);
        const long long c = static_cast<long long>(j + g.ng) * g.nxt + (i + g.ng);
        const double dbx =
            (static_cast<double>(bx[c + 1]) - static_cast<double>(bx[c - 1])) * inv_2dx;
        const double dby =
            (static_cast<double>(by[c + g.nxt]) - static_cast<double>(by[c - g.nxt])) * inv_2dy;
        const double a = __builtin_fabs(dbx + dby);
        m = (m < a) ? a : m;
    }
    m = wave_max(m);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomic_max_speed(out, m);
    }
}

// One wave per component: lane l adds partials l, l+64, ... in order, then a fixed shuffle tree.
template <typename R>
__global__ __launch_bounds__(kNC * kWave) void k_totals_final(const DD* __restrict__ partial,
                                                               int blocks, double* __restrict__ out)
{
    const int c = threadIdx.x >> 6;
    const int lane = threadIdx.x & (kWave - 1);
    DD v{0.0, 0.0};
    for (int b = lane; b < blocks; b += kWave) {
        v = dd_add(v, partial[b * kNC + c]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const DD o{__shfl_down(v.hi, off), __shfl_down(v.lo, off)};
        v = dd_add(v, o);
    }
    if (lane == 0) {
        out[c] = v.hi + v.lo;
    }
}

// Relative L1 distance of state A from a reference state B of the same grid, per component,
// over the interior cells: per block, double-double sums of |a - b| and of |b| (partial[b][c]
// and partial[b][kNC + c]); the final kernel gives sum |a - b| / sum |b| (the absolute sum
// where sum |b| is 0). The difference is taken in fp64 from the stored values, so states of
// either precision compare exactly as stored. Fixed-order reductions: the same bits on every
// call.
template <typename RA, typename RB>
__global__ __launch_bounds__(kTotalsBlock) void k_l1_partial(const RA* __restrict__ A,
                                                             const RB* __restrict__ B, Geom g,
                                                             DD* __restrict__ partial)
{
    __shared__ DD red[2 * kNC][kTotalsBlock / kWave];
    const long long ncell = static_cast<long long>(g.nx) * g.ny;
    DD diff[kNC], ref[kNC];
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        diff[c] = DD{0.0, 0.0};
        ref[c] = DD{0.0, 0.0};
    }
    for (long long k = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; k < ncell;
         k += static_cast<long long>(gridDim.x) * blockDim.x) {
        const int jj = static_cast<int>(k / g.nx);
        const int ii = static_cast<int>(k - static_cast<long long>(jj) * g.nx);
        const long long cell = static_cast<long long>(jj + g.ng) * g.nxt + (ii + g.ng);
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            const double a = static_cast<double>(A[c * g.plane + cell]);
            const double b = static_cast<double>(B[c * g.plane + cell]);
            diff[c] = dd_add(diff[c], fabs(a - b));
            ref[c] = dd_add(ref[c], fabs(b));
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 2 * kNC; ++c) {
        DD v = c < kNC ? diff[c] : ref[c - kNC];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const DD o{__shfl_down(v.hi, off), __shfl_down(v.lo, off)};
            v = dd_add(v, o);
        }
        if (lane == 0) {
            red[c][wave] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2 * kNC) {
        DD v = red[threadIdx.x][0];
        for (int w = 1; w < kTotalsBlock / kWave; ++w) {
            v = dd_add(v, red[threadIdx.x][w]);
        }
        partial[blockIdx.x * 2 * kNC + threadIdx.x] = v;
    }
}

template <int NC>  // kNC; a template so that every translation unit may include it
__global__ __launch_bounds__(NC * kWave) void k_l1_final(const DD* __restrict__ partial,
                                                         int blocks, double* __restrict__ out)
{
    const int c = threadIdx.x >> 6;
    const int lane = threadIdx.x & (kWave - 1);
    DD d{0.0, 0.0}, r{0.0, 0.0};
    for (int b = lane; b < blocks; b += kWave) {
        d = dd_add(d, partial[b * 2 * NC + c]);
        r = dd_add(r, partial[b * 2 * NC + NC + c]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const DD od{__shfl_down(d.hi, off), __shfl_down(d.lo, off)};
        const DD orr{__shfl_down(r.hi, off), __shfl_down(r.lo, off)};
        d = dd_add(d, od);
        r = dd_add(r, orr);
    }
    if (lane == 0) {
        const double sd = d.hi + d.lo;
        const double sr = r.hi + r.lo;
        out[c] = sr > 0.0 ? sd / sr : sd;
    }
}

// ---- fused RK stage ------------------------------------------------------------------------
//
// One wave owns a strip of 60 interior columns and marches rows [r0, r1). Lane l holds global
// column gi = ng + 60 s - 2 + l; lanes 2..61 own cells, lanes 2..62 compute the x-face on the
// left of their column, lanes 0, 1, 63 only supply neighbours. Per row step (written for the
// downward march; the upward one mirrors it, see march_tile):
//   * the new row c+1 is loaded (coalesced, 64 consecutive cells per component) and turned into
//     primitives once;
//   * y: the slope of row c comes from a three-row register window, the face c-1/2 from the
//     carried face state, and the previous y-flux is carried, so each y-face is built once;
//   * x (row c-1): neighbours arrive by lane shuffles, each cell's slope is built once, the
//     left face state is shuffled from lane l-1 and the right flux from lane l+1.
// Ghost cells are either read from memory (after k_ghost_fill) or, with REMAP, addressed at
// their interior source directly, which equals the x-then-y fill.
struct StageBuffers {
    const void* src;  // stage input state (U in stage 1, U* in stage 2)
    const void* u0;   // stage 2: U at the start of the step (aliases dst); stage 1: unused
    void* dst;        // stage 1: U*; stage 2: U (in place)
    const uint8_t* mask;
    Ctrl* ctrl;
    double* dt_log;
    int dt_cap;
};

// Access to the nine component planes of one state array, through a pointer in the global
// address space (global_load / global_store rather than flat accesses, whose completion the
// LDS waits of the march would also have to wait for). An access names a wave-uniform row
// (its address is formed in scalar registers) and a lane's unsigned byte offset from it, so
// that it compiles to a global_load / global_store with a scalar base and a 32-bit vector
// offset: no per-access 64-bit address arithmetic on the vector ALU.
template <typename R>
struct Planes {
    using GR = __attribute__((address_space(1))) R;
    using GB = __attribute__((address_space(1))) char;
    GR* p;
    long long plane;  // elements per plane

    // (q and plane_elems are wave-uniform; they are read into scalar registers here, so that
    // values a caller loaded through vector loads may feed the scalar row addresses below)
    __device__ __forceinline__ static Planes make(const void* q, long long plane_elems)
    {
        const unsigned long long a = uniform_u64(reinterpret_cast<unsigned long long>(q));
        const unsigned long long n = uniform_u64(static_cast<unsigned long long>(plane_elems));
        return Planes{reinterpret_cast<GR*>(a), static_cast<long long>(n)};
    }
    // element `boff` bytes past the start of row r of component c (row length nxt). The row
    // address passes through an empty asm statement that pins it to scalar registers, and the
    // offset through one that keeps it a 32-bit vector register at the access: the compiler
    // can then neither fold the lane's offset into the sum of plane and row offsets (it would
    // form every component's address on the vector ALU) nor widen it to 64 bits away from the
    // access, and the access takes the scalar-base form. (No instruction is emitted for them.)
    __device__ __forceinline__ GR& at(int c, int r, int nxt, unsigned boff) const
    {
        GR* row = p + (c * plane + static_cast<long long>(r) * nxt);
        unsigned long long a = reinterpret_cast<unsigned long long>(row);
        asm("" : "+s"(a));
        asm("" : "+v"(boff));
        return *reinterpret_cast<GR*>(reinterpret_cast<GB*>(a) + boff);
    }
    __device__ __forceinline__ R ld(int c, int r, int nxt, unsigned boff) const
    {
        return at(c, r, nxt, boff);
    }
    __device__ __forceinline__ void st(int c, int r, int nxt, unsigned boff, R v) const
    {
        at(c, r, nxt, boff) = v;
    }
};

// A copy of x in a register of its own (one register move the compiler cannot elide).
__device__ __forceinline__ float own_register(float x)
{
    float y;
    asm("v_mov_b32 %0, %1" : "=v"(y) : "v"(x));
    return y;
}
__device__ __forceinline__ double own_register(double x)
{
    double y;
    asm("v_mov_b64 %0, %1" : "=v"(y) : "v"(x));
    return y;
}

// MUSCL face state w + SIGN d / 2 (include/mhd/kernels/muscl_hll.hpp:59-68). The exact fp64
// build rounds the product and the sum separately, as the CPU path does; the FP32 and the
// contracted fp64 builds take one fused multiply-add per side (the compiler does not fuse the
// product that both sides share by itself).
template <int SIGN, typename R>
__device__ __forceinline__ R face_state(R w, R d)
{
    if constexpr (sizeof(R) == 4) {
        return __builtin_fmaf(SIGN * 0.5f, d, w);
    } else if constexpr (MHD_FAST_FP64) {
        return __builtin_fma(SIGN * 0.5, d, w);
    } else {
        return SIGN > 0 ? w + R(0.5) * d : w - R(0.5) * d;
    }
}

// A cell's right-hand side from its flux differences, -dF (1/dx) - dG (1/dy)
// (include/mhd/rhs.hpp:116-123), and the two RK combines u + dt L and 1/2 (u0 + u) + dt/2 L
// (include/mhd/field_ops.hpp:47-63). The contracted builds take their fused multiply-adds
// from the source: left to the compiler, which of the two products it fused depended on the
// direction the rows were marched in, and the two directions must give the same bits.
template <typename R>
__device__ __forceinline__ R cell_rhs(R dF, R dG, R inv_dx, R inv_dy)
{
    if constexpr (sizeof(R) == 4) {
        return __builtin_fmaf(-dG, inv_dy, -(dF * inv_dx));
    } else if constexpr (MHD_FAST_FP64) {
        return __builtin_fma(-dG, inv_dy, -(dF * inv_dx));
    } else {
        return -dF * inv_dx - dG * inv_dy;
    }
}
template <typename R>
__device__ __forceinline__ R rk_stage1(R u, R dt, R l)
{
    if constexpr (sizeof(R) == 4) {
        return __builtin_fmaf(dt, l, u);
    } else if constexpr (MHD_FAST_FP64) {
        return __builtin_fma(dt, l, u);
    } else {
        return u + dt * l;
    }
}
template <typename R>
__device__ __forceinline__ R rk_stage2(R u0, R u, R half_dt, R l)
{
    const R half(0.5);
    if constexpr (sizeof(R) == 4) {
        return __builtin_fmaf(half_dt, l, half * (u0 + u));
    } else if constexpr (MHD_FAST_FP64) {
        return __builtin_fma(half_dt, l, half * (u0 + u));
    } else {
        return half * (u0 + u) + half_dt * l;
    }
}

// Stage 1 marches its rows in increasing order and stage 2 in decreasing order, so that each
// stage starts on the rows the previous one touched last (still in the memory-side cache).
template <int STAGE>
constexpr int kMarchDir = STAGE == 2 ? -1 : 1;

// March rows [r0, r1) of one 60-column strip (see the comment above StageBuffers) and apply the
// RK stage to its cells. `ring` is this wave's LDS slice. Returns the wave-local maximum CFL
// signal speed of the updated fluid cells (stage 2; 0 in stage 1).
//
// HALF 2: a tile with dB > 0 (wave-uniform) is a narrow strip (at most kHalfCells columns,
// the tail of a grid) and the wave works as two half-wave strips of 32 lanes, each with its
// own 2 halo lanes on either side. Lanes 0..31 march rows [r0, r1), lanes 32..63 rows
// [r0 + dB, r1 + dB), of which only those below r1b are updated. Shifts across lanes 31/32
// feed only halo lanes, whose faces are never used (the same argument as for the wave's edge
// lanes). A tile with dB = 0 is a whole-wave strip. HALF 0: whole-wave strips only.
//
// EARLY: stage 2 issues its loads of U for the updated row at the start of the row step
// (latency overlapped with the faces, nine more live registers); otherwise at the update.
//
// Every row step issues the same loads (the rows are clamped where the march runs out), U
// first and the next row second: the wait before the update then covers only the loads of U,
// and the wait for the next row at the top of the following step lets the stores of the
// update stay in flight.
//
// DIR: the order in which the rows are visited (see the loop below).
template <int STAGE, int LIM, int NI, int REMAP, int HALF, int EARLY, int DIR, typename R,
          typename Planes>
__device__ __forceinline__ R march_tile(const Planes& src, const Planes& u0, const Planes& dst,
                                        const uint8_t* mask, const Geom& g, const Physics& ph,
                                        const StepState& st, int strip, int r0, int r1, int dB,
                                        int r1b, R (&ring)[3][kNC][kWave], int lane)
{
    static_assert(HALF == 0 || HALF == 2, "whole-wave tiles, or both kinds");
    const bool halves = HALF == 2 && dB > 0;  // wave-uniform
    const bool upper = halves && lane >= kWave / 2;
    const int l = halves ? (lane & (kWave / 2 - 1)) : lane;  // lane within its strip
    const int row_off = upper ? dB : 0;                        // this lane's row offset
    // the same in cells; dB <= ny / 2 + 1 keeps it (and the row difference in load_row) below
    // half a plane plus a row, within int for any grid whose two states fit in HBM
    const int lane_off = row_off * g.nxt;
    const int gi = g.ng + strip * kStripCells - 2 + l;
    const int gic = min(gi, g.nxt - 1);
    const int col = REMAP ? ghost_source(gic, g.nx, g.ng, g.bc_x) : gic;
    const bool cell_lane =
        l >= 2 && l < 2 + (halves ? kHalfCells : kStripCells) && gi < g.ng + g.nx;
    // byte offsets of this lane's column in a row it loads, and of its cell in the row it
    // updates (lane_off included), from the start of the row
    const unsigned col_boff = static_cast<unsigned>(col) * static_cast<unsigned>(sizeof(R));
    const unsigned cell_boff =
        static_cast<unsigned>(gic + lane_off) * static_cast<unsigned>(sizeof(R));
    const unsigned long long x_live = halves ? kXFaceLanesHalf : kXFaceLanes;

    const R gamma = static_cast<R>(ph.gamma);
    const R gm1 = static_cast<R>(ph.gm1);
    const R inv_gm1 = static_cast<R>(ph.inv_gm1);
    const R c_h = static_cast<R>(st.c_h);
    const R inv_dx = static_cast<R>(ph.inv_dx);
    const R inv_dy = static_cast<R>(ph.inv_dy);
    const R inv_2dx = static_cast<R>(ph.inv_2dx);
    const R inv_2dy = static_cast<R>(ph.inv_2dy);
    const R half(0.5);
    // include/mhd/rhs.hpp:173-174
    const R psi_damp = (st.c_h > 0.0 && ph.glm_alpha > 0.0)
                           ? static_cast<R>(st.c_h * ph.glm_alpha / ph.h)
                           : R(0);
    const R dt = static_cast<R>(st.dt);
    const R half_dt = static_cast<R>(0.5 * st.dt);

    auto load_row = [&](int q, R (&u)[kNC]) {
        int qbase;
        unsigned boff = col_boff;
        if constexpr (HALF == 2) {
            // both halves' rows are wave-uniform values; the lower one is the base and a lane
            // adds its own row's offset. Without the remap the upper half's last chunk of an
            // odd ny reaches one row past the grid, a row whose values feed only faces of rows
            // that are not updated: it is clamped to the last ghost row.
            const int qa = REMAP ? ghost_source(q, g.ny, g.ng, g.bc_y) : q;
            const int qb = REMAP ? ghost_source(q + dB, g.ny, g.ng, g.bc_y)
                                 : min(q + dB, g.nyt - 1);
            qbase = min(qa, qb);
            boff += static_cast<unsigned>(((upper ? qb : qa) - qbase) * g.nxt) *
                    static_cast<unsigned>(sizeof(R));
        } else {
            qbase = REMAP ? ghost_source(q, g.ny, g.ng, g.bc_y) : q;
        }
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            u[c] = src.ld(c, qbase, g.nxt, boff);
        }
    };

    // The march visits rows c = c_first, c_first + DIR, ..., c_last. Wa, Wb, Wn
    // hold the primitives of rows c - DIR, c, c + DIR: with DIR = 1 (downward in memory order)
    // from c = r0 - 1 to r1, the face c - 1/2 is formed at row c and row c - 1 updated; with
    // DIR = -1 from c = r1 to r0 - 1, the face c + 1/2 and row c + 1. Both form every face from
    // the same operands, so the results are the same bits.
    const int c_first = DIR > 0 ? r0 - 1 : r1;
    const int c_last = DIR > 0 ? r1 : r0 - 1;
    R Wa[kNC], Wb[kNC], Wn[kNC], Uraw[kNC];
    R Wc[kNC], Gp[kNC];  // the carried face state of row c - DIR, the flux of its other face
    R Jp[3] = {R(0), R(0), R(0)};
    R local_speed = R(0);

    // ring slot of row q, counted from the segment start so that it is the same on all lanes
    auto slot_of = [&](int q) { return (q - r0 + 3) % 3; };
    auto ring_put = [&](int q, const R (&u)[kNC]) {
#pragma unroll
        for (int k = 0; k < kNC; ++k) {
            ring[slot_of(q)][k][lane] = u[k];
        }
    };

    load_row(c_first - DIR, Uraw);
    to_prim(Uraw, gm1, Wa);
    load_row(c_first, Uraw);
    to_prim(Uraw, gm1, Wb);
    load_row(c_first + DIR, Uraw);

    // (a row step's place in the march, from c alone: the first step forms no face, the
    // second no update, the last loads no row)
    for (int c = c_first; DIR > 0 ? c <= c_last : c >= c_last; c += DIR) {
        const bool face = DIR > 0 ? c >= r0 : c <= r1 - 1;
        const bool update = DIR > 0 ? c >= r0 + 1 : c <= r1 - 2;
        ring_put(c + DIR, Uraw);
        to_prim(Uraw, gm1, Wn);  // row c + DIR
        // The primitives that equal their conserved values (rho, B, psi) take registers of
        // their own, so that the raw row's registers are free once the row is in the ring and
        // the next row can be loaded into them: otherwise the loaded values are moved into
        // place at the end of the step, and each move waits for its load and for everything
        // issued before it, the update's stores included.
        Wn[RHO] = own_register(Wn[RHO]);
        Wn[BX] = own_register(Wn[BX]);
        Wn[BY] = own_register(Wn[BY]);
        Wn[BZ] = own_register(Wn[BZ]);
        Wn[PSI] = own_register(Wn[PSI]);
        // stage 2 updates row c - DIR from U at the start of the step: with EARLY those loads
        // are issued now, so that their latency overlaps the face computations (column clamped
        // like the row loads; a step with no update loads a row of the tile that it does not
        // use)
        constexpr bool early_u0 = STAGE == 2 && EARLY;
        R U0r[kNC];
        if constexpr (early_u0) {
            const int ru = DIR > 0 ? max(c - DIR, r0) : min(c - DIR, r1 - 1);
#pragma unroll
            for (int k = 0; k < kNC; ++k) {
                U0r[k] = u0.ld(k, ru, g.nxt, cell_boff);
            }
        }
        // row c + 2 DIR for the next step (the last step reloads row c + DIR, unused)
        load_row(DIR > 0 ? min(c + 2, r1 + 1) : max(c - 2, r0 - 2), Uraw);
        // the rows below and above row c
        const R (&Wlo)[kNC] = DIR > 0 ? Wa : Wn;
        const R (&Whi)[kNC] = DIR > 0 ? Wn : Wa;
        // y slope of row c and the two face states it produces: toward row c - DIR (used now)
        // and toward row c + DIR (carried)
        R Wf[kNC], Wn_c[kNC];
#pragma unroll
        for (int k = 0; k < kNC; ++k) {
            const R d = limit_slope<LIM>(Wb[k] - Wlo[k], Whi[k] - Wb[k]);
            const R lo = face_state<-1>(Wb[k], d);  // the state at face c - 1/2
            const R hi = face_state<1>(Wb[k], d);   // the state at face c + 1/2
            Wf[k] = DIR > 0 ? lo : hi;
            Wn_c[k] = DIR > 0 ? hi : lo;
        }
        R Jc[3] = {R(0), R(0), R(0)};
        if (NI) {
            // J = curl B at (gi, c): include/mhd/nonideal.hpp:30-60
            const R byL = from_left(Wb[BY]);
            const R byR = from_right(Wb[BY]);
            const R bzL = from_left(Wb[BZ]);
            const R bzR = from_right(Wb[BZ]);
            const R dbz_dy = (Whi[BZ] - Wlo[BZ]) * inv_2dy;
            const R dbz_dx = (bzR - bzL) * inv_2dx;
            const R dby_dx = (byR - byL) * inv_2dx;
            const R dbx_dy = (Whi[BX] - Wlo[BX]) * inv_2dy;
            Jc[0] = dbz_dy;
            Jc[1] = -dbz_dx;
            Jc[2] = dby_dx - dbx_dy;
        }
        if (face) {
            // y-face between rows c - DIR (Wa) and c (Wb), its lower row's state on the left
            R G[kNC];
            if constexpr (DIR > 0) {
                hll_face<1>(Wc, Wf, Wa, Wb, gamma, inv_gm1, c_h, G, ~0ull);
            } else {
                hll_face<1>(Wf, Wc, Wb, Wa, gamma, inv_gm1, c_h, G, ~0ull);
            }
            if (NI) {
                // include/mhd/rhs.hpp:96-114
                R Jf[3], Bf[3], E[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Jf[k] = half * (Jp[k] + Jc[k]);
                    Bf[k] = half * (Wa[BX + k] + Wb[BX + k]);
                }
                nonideal_e(Jf, Bf, ph.ni, E);
                G[BX] += E[2];
                G[BZ] += -E[0];
                G[PR] += Bf[2] * E[0] - Bf[0] * E[2];
            }
            if (update) {
                // x-direction of row c - DIR (primitives Wa), then assemble and update it
                const int rr = c - DIR;
                R Wl[kNC], WRx[kNC], WLx[kNC];
#pragma unroll
                for (int k = 0; k < kNC; ++k) {
                    Wl[k] = from_left(Wa[k]);
                    const R wr = from_right(Wa[k]);
                    const R d = limit_slope<LIM>(Wa[k] - Wl[k], wr - Wa[k]);
                    WRx[k] = face_state<-1>(Wa[k], d);
                    WLx[k] = from_left(face_state<1>(Wa[k], d));
                }
                R F[kNC];
                hll_face<0>(WLx, WRx, Wl, Wa, gamma, inv_gm1, c_h, F, x_live);
                if (NI) {
                    // include/mhd/rhs.hpp:74-93
                    R Jf[3], Bf[3], E[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        Jf[k] = half * (from_left(Jp[k]) + Jp[k]);
                        Bf[k] = half * (Wl[BX + k] + Wa[BX + k]);
                    }
                    nonideal_e(Jf, Bf, ph.ni, E);
                    F[BY] += -E[2];
                    F[BZ] += E[1];
                    F[PR] += Bf[1] * E[2] - Bf[2] * E[1];
                }
                // the y-fluxes of row rr's upper (rr + 1/2) and lower (rr - 1/2) faces
                const R (&Gup)[kNC] = DIR > 0 ? G : Gp;
                const R (&Glo)[kNC] = DIR > 0 ? Gp : G;
                R L[kNC];
#pragma unroll
                for (int k = 0; k < kNC; ++k) {
                    const R Fr = from_right(F[k]);
                    // include/mhd/rhs.hpp:116-123
                    L[k] = cell_rhs(Fr - F[k], Gup[k] - Glo[k], inv_dx, inv_dy);
                }
                if (psi_damp != R(0)) {
                    L[PSI] -= psi_damp * Wa[PSI];
                }
                const int rl = rr + row_off;  // this lane's row
                if (cell_lane && (!upper || rl < r1b)) {
                    if (mask != nullptr &&
                        mask[static_cast<long long>(rl - g.ng) * g.nx + (gi - g.ng)] == 0) {
#pragma unroll
                        for (int k = 0; k < kNC; ++k) {
                            L[k] = R(0);
                        }
                    }
                    // (a cell lane's column is in the grid, so gic = gi here)
                    const int slot = slot_of(rr);
                    if (STAGE == 1) {
                        // include/mhd/field_ops.hpp:47-53
#pragma unroll
                        for (int k = 0; k < kNC; ++k) {
                            dst.st(k, rr, g.nxt, cell_boff,
                                   rk_stage1(ring[slot][k][lane], dt, L[k]));
                        }
                    } else {
                        // include/mhd/field_ops.hpp:55-63
                        R un[kNC];
#pragma unroll
                        for (int k = 0; k < kNC; ++k) {
                            const R u0k = early_u0 ? U0r[k] : u0.ld(k, rr, g.nxt, cell_boff);
                            un[k] = rk_stage2(u0k, ring[slot][k][lane], half_dt, L[k]);
                        }
#pragma unroll
                        for (int k = 0; k < kNC; ++k) {
                            dst.st(k, rr, g.nxt, cell_boff, un[k]);
                        }
                        const bool fluid =
                            mask == nullptr ||
                            mask[static_cast<long long>(rl - g.ng) * g.nx + (gi - g.ng)] != 0;
                        if (fluid) {
                            local_speed = vmax(local_speed, cell_max_speed(un, gamma, gm1));
                        }
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < kNC; ++k) {
                Gp[k] = G[k];
            }
        }
#pragma unroll
        for (int k = 0; k < kNC; ++k) {
            Wc[k] = Wn_c[k];
            Wa[k] = Wb[k];
            Wb[k] = Wn[k];
        }
        if (NI) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Jp[k] = Jc[k];
            }
        }
    }
    return local_speed;
}

template <int STAGE, int LIM, int NI, int REMAP, typename R, int FAST = 0>
__global__ __launch_bounds__(kBlock) void k_stage(StageBuffers b, Geom g, Physics ph, Tiling tl,
                                                  int parity)
{
    static_assert(FAST == MHD_FAST_FP64, "stepping kernel built in the wrong translation unit");
    const StepState st = step_state(b.ctrl, ph, parity);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (STAGE == 1 && lead) {
        b.ctrl->chmax[parity ^ 1] = 0ull;
        if (!st.active && ctrl_pending(b.ctrl, ph, parity)) {
            b.ctrl->bad_dt = 1;
        }
    }
    if (STAGE == 2 && lead) {
        const int q = parity ^ 1;
        b.ctrl->t[q] = st.active ? st.t + st.dt : st.t;
        b.ctrl->steps[q] = st.active ? st.steps + 1 : st.steps;
        b.ctrl->ch_last[q] = st.active ? st.c_h : b.ctrl->ch_last[parity];
        if (st.active && b.dt_log != nullptr && st.steps < b.dt_cap) {
            b.dt_log[st.steps] = st.dt;
        }
    }
    if (!st.active) {
        return;
    }
    // Tiles in row-band order (a band's strips left to right), four consecutive ones per block:
    // the four waves march the same rows of neighbouring strips at about the same time, so the
    // halo columns and the cache lines shared across their strip edges meet in L2. Consecutive
    // logical blocks are placed on the same XCD, which thus walks down a run of bands.
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_in_block();
    // (stage 2 takes each XCD's blocks in reverse order, so that it starts on the tiles that
    // stage 1 touched last)
    const int logical = xcd_logical_block(blockIdx.x, gridDim.x, STAGE == 2);
    // (the host caps a grid at 2^31 tiles, so 32-bit unsigned arithmetic holds)
    const unsigned wl = static_cast<unsigned>(logical) * kWavesPerBlock + wave;
    const int strip = static_cast<int>(wl % static_cast<unsigned>(tl.strips));
    const int seg = static_cast<int>(wl / static_cast<unsigned>(tl.strips));
    if (seg >= tl.segs) {
        return;
    }
    const int r0 = g.ng + seg * tl.seg_rows;
    const int r1 = min(r0 + tl.seg_rows, g.ng + g.ny);

    const Planes<R> src = Planes<R>::make(b.src, g.plane);
    const Planes<R> u0 = Planes<R>::make(b.u0, g.plane);
    const Planes<R> dst = Planes<R>::make(b.dst, g.plane);
    // wave-private ring of the last three raw (conserved) rows: the RK update of row c-1 reads
    // its stage input from here instead of global memory
    __shared__ R ring[kWavesPerBlock][3][kNC][kWave];
    // Odd row segments march the other way: the two rows each side of a segment boundary are
    // read by both segments' waves, and with the directions alternating both read them at the
    // same end of their march (at about the same time), so the second read finds them in L2.
    const R local_speed =
        (seg & 1) ? march_tile<STAGE, LIM, NI, REMAP, 0, 1, -kMarchDir<STAGE>, R>(
                        src, u0, dst, b.mask, g, ph, st, strip, r0, r1, 0, r1, ring[wave], lane)
                  : march_tile<STAGE, LIM, NI, REMAP, 0, 1, kMarchDir<STAGE>, R>(
                        src, u0, dst, b.mask, g, ph, st, strip, r0, r1, 0, r1, ring[wave], lane);
    if (STAGE == 2) {
        const R m = wave_max(local_speed);
        if (lane == 0) {
            atomic_max_speed(&b.ctrl->chmax[parity ^ 1], static_cast<double>(m));
        }
    }
}

}  // namespace mhdhip
