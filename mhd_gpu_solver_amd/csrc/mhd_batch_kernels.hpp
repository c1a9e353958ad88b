// Batched stage kernel: many independent grids of different sizes advance in one persistent
// launch per RK stage, each with its own dt, gamma, CFL number, boundary types and end time.
//
// The work list is a device table of tiles (instance, strip, row range) grouped by wave: wave
// w of the persistent grid (its logical index after the XCD-aware remap of its workgroup)
// marches tiles[first[w] .. first[w + 1]), a list the host plans so that every wave carries
// about the same work and consecutive waves, which share an XCD, take neighbouring tiles. The
// kernel's arguments are table pointers only: the host may add and remove instances between
// launches (by copies on the same stream) and a captured graph stays valid.
#pragma once

#include "mhd_batch_plan.hpp"
#include "mhd_kernels.hpp"

namespace mhdhip {

static_assert(kPlanStripCells == kStripCells && kPlanHalfCells == kHalfCells,
              "the host planner and the kernels must agree on the strip widths");

struct BatchInst {
    void* U;                // this instance's 9 planes of U
    void* Us;               // and of U*
    const uint8_t* mask;    // interior mask (1 = fluid), null when all cells are fluid
    const long long* plan;  // solid-cell fill plan (dst, src or -1, axis), reflective walls
    Geom g;
    Physics ph;
    int plan_count;
    int enabled;
};

struct BatchTables {
    const BatchInst* insts;
    Ctrl* ctrl;  // one control block per instance slot
    const BatchTile* tiles;
    const int* first;       // this launch's per-wave offsets into tiles (waves + 1)
    const int* fill_count;  // how many slots fill_list holds
    const int* fill_list;   // masked slots, whose ghosts and solid cells are filled per stage
    int dry;  // timing pass: stage 2 writes U* instead of U and no control field is touched
};

// Waves per SIMD each batched stage kernel is compiled for: what its registers allow (fp32
// ideal 128 VGPRs, non-ideal 168; fp64 ideal and non-ideal stage 1 256, non-ideal stage 2 512
// with AGPRs: at 256 it would spill 64-88 bytes per lane). The fp32 ideal kernel needs the
// bound: its march for both tile kinds lands one register above 128. Each stage has its own
// persistent grid, so the two stages of one variant may run at different occupancies.
template <typename R, int NI, int STAGE>
constexpr int kBatchWavesPerEU = sizeof(R) == 4 ? (NI ? 3 : 4) : (NI && STAGE == 2 ? 1 : 2);

// REMAP 1: instances without solid cells; ghosts are addressed at their interior sources.
// REMAP 0: instances with a mask; ghosts (and, with reflective walls, the solid cells) are
// read from memory, filled by k_batch_fill before the stage as include/mhd/
// time_integrator.hpp:46-56 fills before every stage, and the mask zeroes the solid cells'
// right-hand side and keeps them out of the CFL speed.
template <int STAGE, int LIM, int NI, int REMAP, typename R, int FAST = 0>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(kBatchWavesPerEU<R, NI, STAGE>)))
void k_batch_stage(BatchTables tb, int parity)
{
    static_assert(FAST == MHD_FAST_FP64, "stepping kernel built in the wrong translation unit");
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = wave_in_block();
    const int gw = xcd_logical_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
    const int t_begin = tb.first[gw];
    const int n_tiles = tb.first[gw + 1] - t_begin;
    __shared__ R ring[kWavesPerBlock][3][kNC][kWave];
    // stage 2 takes the wave's tiles in reverse order (and marches each upward), so that it
    // starts on what stage 1 touched last, and the next stage 1 on what stage 2 touched last
    for (int j = 0; j < n_tiles; ++j) {
        const int t = STAGE == 2 ? t_begin + n_tiles - 1 - j : t_begin + j;
        const BatchTile tile = tb.tiles[t];
        const BatchInst& in = tb.insts[tile.inst];
        if (!in.enabled) {
            continue;
        }
        Ctrl* ctrl = tb.ctrl + tile.inst;
        const StepState st = step_state(ctrl, in.ph, parity);
        // the instance's first tile keeps its step bookkeeping (as block 0 does in k_stage)
        const bool lead = tile.strip == 0 && tile.r0 == in.g.ng && lane == 0;
        if (STAGE == 1 && lead && !tb.dry) {
            ctrl->chmax[parity ^ 1] = 0ull;
            if (!st.active && ctrl_pending(ctrl, in.ph, parity)) {
                ctrl->bad_dt = 1;
            }
        }
        if (STAGE == 2 && lead && !tb.dry) {
            const int q = parity ^ 1;
            ctrl->t[q] = st.active ? st.t + st.dt : st.t;
            ctrl->steps[q] = st.active ? st.steps + 1 : st.steps;
            ctrl->ch_last[q] = st.active ? st.c_h : ctrl->ch_last[parity];
        }
        if (!st.active) {
            continue;
        }
        // the instance's geometry and the tile's rows in scalar registers (they were read by
        // vector loads): the march forms its row addresses from them
        const Geom g = uniform_geom(in.g);
        const Planes<R> src = Planes<R>::make(STAGE == 1 ? in.U : in.Us, g.plane);
        const Planes<R> u0 = Planes<R>::make(in.U, g.plane);
        const Planes<R> dst = Planes<R>::make(STAGE == 1 || tb.dry ? in.Us : in.U, g.plane);
        const int r0 = uniform_i32(tile.r0);
        const int r1 = uniform_i32(tile.r1);
        const int dB = uniform_i32(tile.r0b < 0 ? 0 : tile.r0b - tile.r0);
        const int r1b = uniform_i32(tile.r0b < 0 ? tile.r1 : tile.r1b);
        const int strip = uniform_i32(tile.strip);
        const R local = march_tile<STAGE, LIM, NI, REMAP, 2, 1, kMarchDir<STAGE>, R>(
            src, u0, dst, REMAP ? nullptr : in.mask, g, in.ph, st, strip, r0, r1, dB, r1b,
            ring[wave], lane);
        if (STAGE == 2 && !tb.dry) {
            const R m = wave_max(local);
            if (lane == 0) {
                atomic_max_speed(&ctrl->chmax[parity ^ 1], static_cast<double>(m));
            }
        }
    }
}

// Ghost fill, then solid-wall fill, of the stage input of every listed instance whose step
// with this parity is still pending: one workgroup per instance, so the ghosts take the solid
// cells' values before the wall fill replaces them, as the separate fills do in order
// (include/boundary_conditions/boundary_conditions.hpp:33-103, include/mhd/mask.hpp:85-149).
template <typename R>
__global__ __launch_bounds__(256) void k_batch_fill(BatchTables tb, int parity, int stage)
{
    if (static_cast<int>(blockIdx.x) >= *tb.fill_count) {
        return;
    }
    const int si = tb.fill_list[blockIdx.x];
    const BatchInst& in = tb.insts[si];
    if (!in.enabled || !ctrl_pending(tb.ctrl + si, in.ph, parity)) {
        return;
    }
    R* U = static_cast<R*>(stage == 1 ? in.U : in.Us);
    const long long total = ghost_cells(in.g);
    for (long long t = threadIdx.x; t < total; t += blockDim.x) {
        ghost_fill_cell(U, in.g, t);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < in.plan_count; k += blockDim.x) {
        wall_fill_cell(U, in.g, in.plan, k, in.ph.gamma);
    }
}

}  // namespace mhdhip
