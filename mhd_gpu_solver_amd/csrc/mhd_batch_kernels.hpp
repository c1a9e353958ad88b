// Batched stage kernel: many independent grids of different sizes advance in one persistent
// launch per RK stage, each with its own dt, gamma, CFL number, boundary types and end time.
//
// The work list is a device table of tiles (instance, strip, row range), planned by the host.
// Queued hand-out (the default): one queue of tiles per XCD, which the workgroups on that XCD
// take four neighbouring tiles at a time (see the loop below). Static hand-out (the capacity
// fill): wave w of the persistent grid (its logical index after the XCD-aware remap of its
// workgroup) marches tiles[first[w] .. first[w + 1]). The kernel's arguments are table
// pointers only: the host may add and remove instances between launches (by copies on the
// same stream) and a captured graph stays valid.
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
    int stepping;  // its tiles are in the current tile lists
};

struct BatchTables {
    const BatchInst* insts;
    Ctrl* ctrl;  // one control block per instance slot
    const BatchTile* tiles;
    const int* first;       // static hand-out: this launch's per-wave offsets into tiles
    const int* fill_count;  // how many slots fill_list holds
    const int* fill_list;   // masked slots, whose ghosts and solid cells are filled per stage
    int dry;  // timing pass: stage 2 writes U* instead of U and no control field is touched
    // queued hand-out (null qbegin: the static one): this launch's queue offsets into tiles
    // (launch_queues(grid) + 1), its queue counters, and the counters of the same variant's
    // other stage, which this launch zeroes as it starts
    const int* qbegin;
    int* next;
    int* other_next;
    // Instances of the tile lists still to step, per step parity: live[p] is the count for
    // the step with parity p, constant while its launches run. Stage 1 copies it to
    // live[p ^ 1], stage 2 takes an instance off live[p ^ 1] when it finishes; the host sets
    // live[0] before each launch sequence. A launch with none returns at once (after the
    // bookkeeping an inactive step does) instead of visiting every tile of finished instances.
    int* live;
    int slots;  // instance slots (entries of insts and ctrl)
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
    const bool queued = tb.qbegin != nullptr;
    // static hand-out: this wave's list tiles[t_begin .. t_begin + n_tiles)
    int t_begin = 0, n_tiles = 0;
    // queued hand-out: this wave's queue q (its XCD's) of qn tiles from tiles[qb], which wq
    // waves take, this one first the tile at its index lw among them, then tickets
    int q = 0, qb = 0, qn = 0, wq = 0, lw = 0;
    if (queued) {
        const int nb = static_cast<int>(gridDim.x);
        const int b = static_cast<int>(blockIdx.x);
        q = b % launch_queues(nb);
        qb = uniform_i32(tb.qbegin[q]);
        qn = uniform_i32(tb.qbegin[q + 1]) - qb;
        wq = queue_waves(nb, q, kWavesPerBlock);
        lw = queue_wave_index(nb, b, wave, kWavesPerBlock);
        if (b == 0 && wave == 0 && lane < kPlanXcds) {
            atomicExch(&tb.other_next[lane], 0);
        }
    }
    const int live = uniform_i32(tb.live[parity]);
    if (STAGE == 1 && !tb.dry && blockIdx.x == 0 && threadIdx.x == 0) {
        tb.live[parity ^ 1] = live;
    }
    if (live <= 0) {
        // no instance of the lists steps: the step bookkeeping of each of them, as its lead
        // tile does for an inactive step (none is pending, so none has dt <= 0 to report)
        if (!tb.dry && blockIdx.x == 0) {
            for (int si = static_cast<int>(threadIdx.x); si < tb.slots; si += kBlock) {
                const BatchInst& in = tb.insts[si];
                if (!in.enabled || !in.stepping) {
                    continue;
                }
                Ctrl* ctrl = tb.ctrl + si;
                const int q = parity ^ 1;
                if (STAGE == 1) {
                    ctrl->chmax[q] = 0ull;
                } else {
                    ctrl->t[q] = ctrl->t[parity];
                    ctrl->steps[q] = ctrl->steps[parity];
                    ctrl->ch_last[q] = ctrl->ch_last[parity];
                }
            }
        }
        return;
    }
    if (!queued) {
        const int gw = xcd_logical_block(blockIdx.x, gridDim.x) * kWavesPerBlock + wave;
        t_begin = tb.first[gw];
        n_tiles = tb.first[gw + 1] - t_begin;
    }
    __shared__ R ring[kWavesPerBlock][3][kNC][kWave];
    // Stage 2: the wave holds the largest signal speed of its consecutive tiles of one
    // instance (as the bits of a non-negative double, the order atomic_max_speed compares)
    // and adds it to the instance's control block when the instance changes or the list
    // ends. One atomic per run of tiles instead of one per tile: the atomics of all the tiles
    // of a large grid meet at one address and are served one after another, and each wave's
    // next tile waited for its last one.
    int held = -1;  // instance whose speed the wave holds, -1 for none
    unsigned long long held_bits = 0ull;
    auto release = [&]() {
        if (STAGE == 2 && held >= 0 && lane == 0) {
            atomicMax(&tb.ctrl[held].chmax[parity ^ 1], held_bits);
        }
    };
    // Queued: the waves of a workgroup take neighbouring tiles together (a band's strips come
    // one after another in the queue), first the ones at their own indices, then one ticket of
    // kWavesPerBlock tiles per round, drawn by wave 0 and passed on through LDS (two slots,
    // by round parity, so a slot is rewritten only after every wave has read it). The four
    // waves then march the same rows of neighbouring strips at about the same time, so the
    // cache lines their strips share (halo lanes, the partial lines at strip edges) meet in
    // L2 instead of being fetched and written twice.
    __shared__ int block_ticket[2];
    // stage 2 takes the wave's tiles (the queue's) in reverse order (and marches each upward),
    // so that it starts on what stage 1 touched last, and the next stage 1 on what stage 2
    // touched last
    for (int j = 0;; ++j) {
        int t;
        if (queued) {
            int u = lw;
            if (j > 0) {
                if (wave == 0 && lane == 0) {
                    block_ticket[j & 1] = atomicAdd(&tb.next[q], kWavesPerBlock);
                }
                __syncthreads();
                const int base = wq + uniform_i32(block_ticket[j & 1]);
                if (base >= qn) {
                    break;  // the same for every wave of the workgroup
                }
                u = base + wave;
            }
            if (u >= qn) {
                continue;  // the queue ends inside this round: the next one ends the loop
            }
            t = qb + (STAGE == 2 ? qn - 1 - u : u);
        } else {
            if (j >= n_tiles) {
                break;
            }
            t = STAGE == 2 ? t_begin + n_tiles - 1 - j : t_begin + j;
        }
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
            const double t_new = st.active ? st.t + st.dt : st.t;
            const int steps_new = st.active ? st.steps + 1 : st.steps;
            ctrl->t[q] = t_new;
            ctrl->steps[q] = steps_new;
            ctrl->ch_last[q] = st.active ? st.c_h : ctrl->ch_last[parity];
            if (st.active && !(t_new < in.ph.t_end && steps_new < in.ph.max_steps)) {
                atomicSub(&tb.live[q], 1);  // this instance has finished
            }
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
        // odd bands march the other way (see plan_queues) in the fp64 kernels; the FP32 kernel
        // marches every tile one way: with both directions inlined its stage 2 spills
        // registers and ran 1.5 % slower on the ensemble
        const R local =
            (sizeof(R) == 8 && uniform_i32(tile.flip))
                ? march_tile<STAGE, LIM, NI, REMAP, 2, 1, -kMarchDir<STAGE>, R>(
                      src, u0, dst, REMAP ? nullptr : in.mask, g, in.ph, st, strip, r0, r1, dB,
                      r1b, ring[wave], lane)
                : march_tile<STAGE, LIM, NI, REMAP, 2, 1, kMarchDir<STAGE>, R>(
                      src, u0, dst, REMAP ? nullptr : in.mask, g, in.ph, st, strip, r0, r1, dB,
                      r1b, ring[wave], lane);
        if (STAGE == 2 && !tb.dry) {
            const int inst = uniform_i32(tile.inst);
            if (inst != held) {
                release();
                held = inst;
                held_bits = 0ull;
            }
            const double m = static_cast<double>(wave_max(local));
            const unsigned long long bits =
                static_cast<unsigned long long>(__double_as_longlong(m));
            held_bits = bits > held_bits ? bits : held_bits;
        }
    }
    release();
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
