// Batched ensemble engine and its C ABI (include/mhd_hip.h, mhd_hip_batch_*).
//
// Many independent solves of different grid sizes share one device: each instance keeps its
// state (U and U*) resident in an HBM arena sized by the caller's budget, its own control block
// (t, steps, dt, c_h on the device) and its own physics (gamma, CFL number, end time, boundary
// types). One persistent kernel per RK stage advances every active instance; a hipGraph of
// kSteps steps is replayed and the host reads all control blocks once per replay. Instances
// join (admission, first CFL reduction) and leave between replays by copies on the engine's
// stream, so the captured graph stays valid across changes of the instance set.
#include "mhd_abi.hpp"
#include "mhd_batch_plan.hpp"
#include "mhd_hip.h"
#include "mhd_host.hpp"
#include "mhd_launch.hpp"
#include "mhd_transfer.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

using mhdhip::BatchInst;
using mhdhip::BatchTables;
using mhdhip::BatchTile;
using mhdhip::check;
using mhdhip::Ctrl;
using mhdhip::Full;
using mhdhip::Gate;
using mhdhip::Geom;
using mhdhip::guarded;
using mhdhip::InvalidArgument;
using mhdhip::NoDevice;
using mhdhip::NonIdeal;
using mhdhip::Physics;

namespace {

constexpr size_t kAlign = 256;

size_t round_up(size_t v, size_t a)
{
    return (v + a - 1) / a * a;
}

// First-fit allocator over one device arena (offsets only; the arena is one hipMalloc).
class Arena {
  public:
    void reset(size_t cap)
    {
        cap_ = cap;
        free_.clear();
        if (cap > 0) {
            free_[0] = cap;
        }
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t& off)
    {
        bytes = round_up(bytes, kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest > 0) {
                    free_[off + bytes] = rest;
                }
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = round_up(bytes, kAlign);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t capacity() const { return cap_; }

  private:
    size_t cap_ = 0;
    size_t used_ = 0;
    std::map<size_t, size_t> free_;
};

struct Slot {
    bool used = false;
    int id = -1;
    size_t off = 0;          // arena offset of U; U*, the mask and the wall plan follow
    size_t bytes = 0;        // everything the instance holds in the arena
    size_t state_bytes = 0;  // one state (U or U*), rounded to the arena alignment
    mhd_hip_instance desc{};
    int limiter = MHD_HIP_MC;
    NonIdeal ni{};           // the instance's non-ideal terms
    bool masked = false;     // its mask has solid cells (the mask follows U*)
    bool walls = true;       // solid cells take the reflective wall fill
    int plan_count = 0;      // solid cells in the wall plan (follows the mask)
    bool finished = false;
    bool paused = false;  // resident but not stepped (scheduler's choice)
    mhd_hip_result result{};
    int bad_dt = 0;
};

size_t state_elems(int nx, int ny)
{
    return static_cast<size_t>(nx + 4) * static_cast<size_t>(ny + 4);
}

// include/physics_config/mhd_config.hpp:90-94
bool nonideal_any(const NonIdeal& n)
{
    return (n.ohmic && n.eta_ohm != 0.0) || (n.hall && n.eta_hall != 0.0) ||
           (n.ambipolar && n.eta_ambipolar != 0.0);
}

NonIdeal nonideal_of(const mhd_hip_params& p)
{
    NonIdeal n{};
    n.ohmic = p.ohmic;
    n.hall = p.hall;
    n.ambipolar = p.ambipolar;
    n.eta_ohm = p.eta_ohm;
    n.eta_hall = p.eta_hall;
    n.eta_ambipolar = p.eta_ambipolar;
    return n;
}

// Kernel variant of an instance: limiter (bit 0), non-ideal terms (bit 1), a mask (bit 2:
// ghosts and, with reflective walls, solid cells are filled in memory before each stage
// instead of addressed at their sources).
constexpr int kGroups = 8;
int group_of(const Slot& s)
{
    return (s.masked ? 4 : 0) + (nonideal_any(s.ni) ? 2 : 0) + (s.limiter == MHD_HIP_MC ? 1 : 0);
}
int group_limiter(int g) { return (g & 1) ? MHD_HIP_MC : MHD_HIP_MINMOD; }
bool group_nonideal(int g) { return (g & 2) != 0; }
bool group_remap(int g) { return (g & 4) == 0; }

}  // namespace

struct mhd_hip_batch {
    static constexpr int kSteps = 8;

    int device = -1;
    int precision = MHD_HIP_FP64;
    size_t elem = 8;
    mhd_hip_params phys_template{};  // limiter and non-ideal terms of mhd_hip_batch_add

    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* arena_base = nullptr;
    Arena arena;
    int max_instances = 0;
    std::vector<Slot> slots;
    int next_id = 0;

    BatchInst* d_insts = nullptr;
    Ctrl* d_ctrl = nullptr;
    BatchTile* d_tiles = nullptr;
    int* d_first = nullptr;  // per variant and stage, per-wave tile offsets (max_waves + 1 each)
    int* d_fill = nullptr;   // count, then the slots filled before each stage
    // queued hand-out: per variant and stage, the queue offsets into the tiles (kQueueSlots
    // each), and the queue counters (kPlanXcds each)
    static constexpr int kQueueSlots = mhdhip::kPlanXcds + 1;
    int* d_queue = nullptr;
    int* d_next = nullptr;
    int* d_live = nullptr;  // instances of the tile lists still to step, per step parity
                            // (BatchTables::live)
    size_t tile_cap = 0;

    BatchInst* h_insts = nullptr;  // pinned staging
    Ctrl* h_ctrl = nullptr;
    BatchTile* h_tiles = nullptr;
    int* h_first = nullptr;
    int* h_fill = nullptr;
    int* h_queue = nullptr;
    int* h_live = nullptr;
    std::vector<char> tabled;  // slots whose tiles the current tables hold
    size_t h_tile_cap = 0;
    double* h_stage = nullptr;  // pinned staging for state upload/download
    size_t h_stage_bytes = 0;

    mhdhip::DD* totals_partial = nullptr;
    double* totals_out = nullptr;
    mhdhip::DD* l1_partial = nullptr;  // state-comparison scratch, allocated on first use

    bool dirty = true;
    int sched_policy = MHD_HIP_SCHED_ALL;
    long long sched_cap = 0;  // interior cells stepping per replay under SRPT
    // Work partition: > 0 fixed segments of that many rows, handed out round-robin; 0 the load
    // model's length at every table build; -1 timed on dry steps over a shortlist, then kept;
    // < -1 the capacity fill with pieces of at least -seg_rows rows.
    int seg_rows = 0;
    int tuned_rows = 0;  // the timed choice: > 0 fixed rows, < 0 fill with -tuned_rows min rows
    bool queued = false;  // the current tables hand tiles out from per-XCD queues
    // The queue counters are zeroed by the stage kernels themselves (each stage zeroes the
    // other's as it starts), which holds while every launch of stage 1 is followed by one of
    // stage 2. A launch sequence cut short by an error leaves them unknown: the host zeroes
    // them before the next launches.
    bool counters_known = true;
    int dry_run = 0;     // set while the timed choice launches its dry steps
    int last_rows = 0;  // fixed rows of the current tile list (0: capacity fill)
    int ntiles = 0;
    int blocks = 0;                      // the largest persistent grid of the variants
    int group_blocks[kGroups][2] = {};   // persistent grid of each kernel variant, per stage
    int max_waves = 0;
    unsigned used_groups = 0;            // variants with stepping instances (bit g)
    bool fill_needed = false;            // stepping instances with a mask
    double last_seconds = 0.0;
    hipGraphExec_t graph = nullptr;
    unsigned graph_key = 0;              // the variants and fill the captured graph launches

    // kernel timing mode: replays are launched eagerly with events around each stage kernel
    bool kernel_timing = false;
    double stage_ms[2] = {0.0, 0.0};
    long long stage_launches[2] = {0, 0};
    long long timed_cell_steps = 0;  // interior cells x steps of the active instances timed
    std::vector<hipEvent_t> tev;

    ~mhd_hip_batch()
    {
        if (device >= 0) (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        if (graph) (void)hipGraphExecDestroy(graph);
        if (arena_base) (void)hipFree(arena_base);
        if (d_insts) (void)hipFree(d_insts);
        if (d_ctrl) (void)hipFree(d_ctrl);
        if (d_tiles) (void)hipFree(d_tiles);
        if (d_first) (void)hipFree(d_first);
        if (d_fill) (void)hipFree(d_fill);
        if (d_queue) (void)hipFree(d_queue);
        if (d_next) (void)hipFree(d_next);
        if (d_live) (void)hipFree(d_live);
        if (h_insts) (void)hipHostFree(h_insts);
        if (h_ctrl) (void)hipHostFree(h_ctrl);
        if (h_tiles) (void)hipHostFree(h_tiles);
        if (h_first) (void)hipHostFree(h_first);
        if (h_fill) (void)hipHostFree(h_fill);
        if (h_queue) (void)hipHostFree(h_queue);
        if (h_live) (void)hipHostFree(h_live);
        if (h_stage) (void)hipHostFree(h_stage);
        if (totals_partial) (void)hipFree(totals_partial);
        if (totals_out) (void)hipFree(totals_out);
        if (l1_partial) (void)hipFree(l1_partial);
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // One replay's worth of steps, launched eagerly with an event pair around each stage
    // kernel. After the synchronisation the control blocks tell how many of the kSteps steps
    // stepped some instance (the later launches found every instance finished and returned at
    // once): only those launches and their times are added.
    template <typename R, int FAST>
    void timed_replay()
    {
        std::vector<int> before(static_cast<size_t>(max_instances), 0);
        std::vector<char> stepping(static_cast<size_t>(max_instances), 0);
        for (int si = 0; si < max_instances; ++si) {
            const Slot& s = slots[si];
            stepping[si] = s.used && !s.finished && !s.paused;
            before[si] = h_ctrl[si].steps[0];
        }
        const size_t need = 4 * kSteps;
        while (tev.size() < need) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            tev.push_back(e);
        }
        open_launches();
        for (int k = 0; k < kSteps; ++k) {
            const int p = k & 1;
            check(hipEventRecord(tev[4 * k], stream), "hipEventRecord");
            launch_stage<R, FAST>(1, p);
            check(hipEventRecord(tev[4 * k + 1], stream), "hipEventRecord");
            check(hipEventRecord(tev[4 * k + 2], stream), "hipEventRecord");
            launch_stage<R, FAST>(2, p);
            check(hipEventRecord(tev[4 * k + 3], stream), "hipEventRecord");
        }
        counters_known = true;
        check(hipMemcpyAsync(h_ctrl, d_ctrl, sizeof(Ctrl) * max_instances, hipMemcpyDeviceToHost,
                             stream),
              "hipMemcpyAsync(ctrl)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::vector<int> after(static_cast<size_t>(max_instances), 0);
        for (int si = 0; si < max_instances; ++si) {
            after[si] = h_ctrl[si].steps[0];
        }
        const int live = mhdhip::replay_active_launches(before, after, stepping, kSteps);
        for (int k = 0; k < live; ++k) {
            float a = 0.0f, b = 0.0f;
            check(hipEventElapsedTime(&a, tev[4 * k], tev[4 * k + 1]), "hipEventElapsedTime");
            check(hipEventElapsedTime(&b, tev[4 * k + 2], tev[4 * k + 3]), "hipEventElapsedTime");
            stage_ms[0] += a;
            stage_ms[1] += b;
        }
        stage_launches[0] += live;
        stage_launches[1] += live;
    }

    void init(int prec, const mhd_hip_params& tmpl, size_t budget, int max_inst)
    {
        if (prec != MHD_HIP_FP64 && prec != MHD_HIP_FP32 && prec != MHD_HIP_FP64_FAST) {
            throw InvalidArgument(
                "precision must be MHD_HIP_FP64, MHD_HIP_FP32 or MHD_HIP_FP64_FAST");
        }
        if (tmpl.limiter != MHD_HIP_MINMOD && tmpl.limiter != MHD_HIP_MC) {
            throw InvalidArgument("unknown slope limiter");
        }
        if (max_inst < 1) {
            throw InvalidArgument("max_instances must be positive");
        }
        if (budget < kAlign) {
            throw InvalidArgument("HBM budget too small");
        }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
            throw NoDevice("no HIP device available");
        }
        check(hipGetDevice(&device), "hipGetDevice");
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            throw NoDevice("current HIP device is not gfx950 (MI355X)");
        }
        precision = prec;
        elem = prec == MHD_HIP_FP32 ? 4 : 8;
        phys_template = tmpl;
        max_instances = max_inst;
        slots.assign(max_inst, Slot{});
        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&ev0), "hipEventCreate");
        check(hipEventCreate(&ev1), "hipEventCreate");
        budget = round_up(budget, kAlign);
        check(hipMalloc(&arena_base, budget), "hipMalloc(arena)");
        arena.reset(budget);
        check(hipMalloc(reinterpret_cast<void**>(&d_insts), sizeof(BatchInst) * max_inst),
              "hipMalloc(insts)");
        check(hipMalloc(reinterpret_cast<void**>(&d_ctrl), sizeof(Ctrl) * max_inst),
              "hipMalloc(ctrl)");
        for (int g = 0; g < kGroups; ++g) {
            const int lim = group_limiter(g);
            const bool ni = group_nonideal(g), remap = group_remap(g);
            for (int st = 1; st <= 2; ++st) {
                const int per_cu =
                    precision == MHD_HIP_FP64
                        ? mhdhip::batch_blocks_per_cu<double, 0>(st, lim, ni, remap)
                        : precision == MHD_HIP_FP64_FAST
                              ? mhdhip::batch_blocks_per_cu<double, 1>(st, lim, ni, remap)
                              : mhdhip::batch_blocks_per_cu<float, 0>(st, lim, ni, remap);
                group_blocks[g][st - 1] = std::max(1, per_cu) * prop.multiProcessorCount;
                blocks = std::max(blocks, group_blocks[g][st - 1]);
            }
        }
        max_waves = blocks * mhdhip::kWavesPerBlock;
        const size_t nfirst = static_cast<size_t>(kGroups) * 2 * (max_waves + 1);
        check(hipMalloc(reinterpret_cast<void**>(&d_first), sizeof(int) * nfirst),
              "hipMalloc(first)");
        check(hipMalloc(reinterpret_cast<void**>(&d_fill), sizeof(int) * (max_inst + 1)),
              "hipMalloc(fill)");
        const size_t nqueue = static_cast<size_t>(kGroups) * 2 * kQueueSlots;
        const size_t nnext = static_cast<size_t>(kGroups) * 2 * mhdhip::kPlanXcds;
        check(hipMalloc(reinterpret_cast<void**>(&d_queue), sizeof(int) * nqueue),
              "hipMalloc(queue)");
        check(hipMalloc(reinterpret_cast<void**>(&d_next), sizeof(int) * nnext),
              "hipMalloc(next)");
        check(hipHostMalloc(reinterpret_cast<void**>(&h_queue), sizeof(int) * nqueue, 0),
              "hipHostMalloc");
        std::fill(h_queue, h_queue + nqueue, 0);
        check(hipMemcpyAsync(d_queue, h_queue, sizeof(int) * nqueue, hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync");
        check(hipMemsetAsync(d_next, 0, sizeof(int) * nnext, stream), "hipMemsetAsync(next)");
        check(hipMalloc(reinterpret_cast<void**>(&d_live), 2 * sizeof(int)), "hipMalloc(live)");
        check(hipHostMalloc(reinterpret_cast<void**>(&h_live), 2 * sizeof(int), 0),
              "hipHostMalloc");
        h_live[0] = h_live[1] = 0;
        check(hipMemsetAsync(d_live, 0, 2 * sizeof(int), stream), "hipMemsetAsync(live)");
        tabled.assign(static_cast<size_t>(max_inst), 0);
        check(hipHostMalloc(reinterpret_cast<void**>(&h_insts), sizeof(BatchInst) * max_inst, 0),
              "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&h_ctrl), sizeof(Ctrl) * max_inst, 0),
              "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&h_first), sizeof(int) * nfirst, 0),
              "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&h_fill), sizeof(int) * (max_inst + 1), 0),
              "hipHostMalloc");
        std::memset(h_insts, 0, sizeof(BatchInst) * max_inst);
        std::memset(h_ctrl, 0, sizeof(Ctrl) * max_inst);
        std::fill(h_first, h_first + nfirst, 0);
        std::fill(h_fill, h_fill + max_inst + 1, 0);
        check(hipMemcpyAsync(d_insts, h_insts, sizeof(BatchInst) * max_inst,
                             hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync");
        check(hipMemcpyAsync(d_first, h_first, sizeof(int) * nfirst, hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync");
        check(hipMemcpyAsync(d_fill, h_fill, sizeof(int) * (max_inst + 1), hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync");
        ensure_tiles(1024);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void ensure_tiles(size_t n)
    {
        if (n <= tile_cap) {
            return;
        }
        const size_t cap = std::max(n, tile_cap * 2);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (d_tiles) (void)hipFree(d_tiles);
        if (h_tiles) (void)hipHostFree(h_tiles);
        check(hipMalloc(reinterpret_cast<void**>(&d_tiles), sizeof(BatchTile) * cap),
              "hipMalloc(tiles)");
        check(hipHostMalloc(reinterpret_cast<void**>(&h_tiles), sizeof(BatchTile) * cap, 0),
              "hipHostMalloc(tiles)");
        tile_cap = cap;
        h_tile_cap = cap;
        if (graph) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        dirty = true;
    }

    // Pinned staging of at least `bytes`. A larger buffer replaces the old one only after the
    // stream has drained, so no copy queued from or into the old buffer is still in flight.
    void ensure_stage(size_t bytes)
    {
        if (bytes <= h_stage_bytes) {
            return;
        }
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (h_stage) (void)hipHostFree(h_stage);
        h_stage = nullptr;
        check(hipHostMalloc(reinterpret_cast<void**>(&h_stage), bytes, 0), "hipHostMalloc(stage)");
        h_stage_bytes = bytes;
    }

    static size_t state_bytes_of(int prec, int nx, int ny)
    {
        const size_t e = prec == MHD_HIP_FP32 ? 4 : 8;
        return round_up(9 * state_elems(nx, ny) * e, kAlign);
    }

    // Arena bytes of an instance: U and U*, and with solid cells the mask and the wall plan.
    static size_t instance_bytes(int prec, int nx, int ny, bool masked = false,
                                 size_t plan_count = 0)
    {
        size_t b = 2 * state_bytes_of(prec, nx, ny);
        if (masked) {
            b += round_up(static_cast<size_t>(nx) * ny, kAlign) +
                 round_up(3 * sizeof(long long) * plan_count, kAlign);
        }
        return b;
    }

    Slot& slot_of(int id)
    {
        for (auto& s : slots) {
            if (s.used && s.id == id) {
                return s;
            }
        }
        throw InvalidArgument("unknown instance id " + std::to_string(id));
    }

    int slot_index(const Slot& s) const { return static_cast<int>(&s - slots.data()); }

    char* base_of(const Slot& s) const { return static_cast<char*>(arena_base) + s.off; }
    const uint8_t* mask_of(const Slot& s) const
    {
        return s.masked ? reinterpret_cast<const uint8_t*>(base_of(s) + 2 * s.state_bytes)
                        : nullptr;
    }
    const long long* plan_of(const Slot& s) const
    {
        if (!s.masked) return nullptr;
        const size_t mask_bytes = round_up(static_cast<size_t>(s.desc.nx) * s.desc.ny, kAlign);
        return reinterpret_cast<const long long*>(base_of(s) + 2 * s.state_bytes + mask_bytes);
    }

    // The instance's Physics: gamma, CFL number, spacings, end time and step limit from its
    // description, the non-ideal terms and their diffusive dt bound from its own settings
    // (include/physics_config/mhd_config.hpp:96-110, include/mhd/cfl.hpp:77-89).
    static Physics physics_of(const Slot& s)
    {
        const mhd_hip_instance& d = s.desc;
        const double dx = (d.x_max - d.x_min) / static_cast<double>(d.nx);
        const double dy = (d.y_max - d.y_min) / static_cast<double>(d.ny);
        Physics ph{};
        ph.gamma = d.gamma;
        ph.gm1 = d.gamma - 1.0;
        ph.inv_gm1 = 1.0 / (d.gamma - 1.0);
        ph.cfl = d.cfl;
        ph.inv_dx = 1.0 / dx;
        ph.inv_dy = 1.0 / dy;
        ph.inv_2dx = 0.5 / dx;
        ph.inv_2dy = 0.5 / dy;
        ph.glm_alpha = d.glm_alpha;
        ph.h = (dy < dx) ? dy : dx;
        const NonIdeal& t = s.ni;
        double eta = 0.0;
        auto mx = [](double a, double b) { return (a < b) ? b : a; };
        if (t.ohmic) eta = mx(eta, std::fabs(t.eta_ohm));
        if (t.hall) eta = mx(eta, std::fabs(t.eta_hall));
        if (t.ambipolar) eta = mx(eta, std::fabs(t.eta_ambipolar));
        ph.has_diff = eta > 0.0 ? 1 : 0;
        ph.dt_diff = ph.has_diff ? d.cfl * ph.h * ph.h / (2.0 * 2.0 * eta) : 0.0;
        ph.t_end = d.t_end;
        ph.max_steps = d.max_steps;
        ph.ni = t;
        return ph;
    }

    static Geom geom_of(const mhd_hip_instance& d)
    {
        Geom g{};
        g.nx = d.nx;
        g.ny = d.ny;
        g.ng = 2;
        g.nxt = d.nx + 4;
        g.nyt = d.ny + 4;
        g.plane = static_cast<long long>(g.nxt) * g.nyt;
        g.bc_x = d.bc_x;
        g.bc_y = d.bc_y;
        return g;
    }

    template <typename R>
    void upload_state(const Slot& s, const double* const comps[9])
    {
        const size_t n = state_elems(s.desc.nx, s.desc.ny);
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");  // staging reuse
        ensure_stage(9 * n * sizeof(R));
        R* st = reinterpret_cast<R*>(h_stage);
        mhdhip::gather_planes(st, comps, 9, n);
        check(hipMemcpyAsync(base_of(s), st, 9 * n * sizeof(R), hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync(upload)");
    }

    template <typename R, int FAST>
    void start_instance(int si)
    {
        const Slot& s = slots[si];
        const Geom g = geom_of(s.desc);
        const Physics ph = physics_of(s);
        R* u = reinterpret_cast<R*>(base_of(s));
        check(mhdhip::launch_ctrl_init<R>(d_ctrl + si, stream), "ctrl_init");
        // include/mhd/cfl.hpp:46-71: the fluid interior cells only
        check(mhdhip::launch_cfl<R, FAST>(u, g, ph, mask_of(s), d_ctrl + si, stream), "cfl");
        // the host copy of this slot's control block, so its progress reads right before the
        // first replay
        check(hipMemcpyAsync(h_ctrl + si, d_ctrl + si, sizeof(Ctrl), hipMemcpyDeviceToHost,
                             stream),
              "hipMemcpyAsync(ctrl)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // Admission of one instance with its own limiter and non-ideal terms (`scheme`), an
    // optional interior mask (null: all fluid; include/mhd/mask.hpp:17-83) and the reflective
    // wall fill of its solid cells (walls; include/mhd/mask.hpp:85-149).
    int add(const mhd_hip_instance& d, const mhd_hip_params& scheme, const uint8_t* mask,
            size_t mask_len, bool walls, const double* const comps[9])
    {
        if (d.nx < 1 || d.ny < 1) {
            throw InvalidArgument("instance grid must have at least one interior cell per axis");
        }
        if (d.t_end < 0.0) {
            throw InvalidArgument("t_end must be non-negative");
        }
        if (d.cfl <= 0.0) {
            throw InvalidArgument("CFL number must be positive");
        }
        if (d.max_steps < 0) {
            throw InvalidArgument("max_steps must be non-negative");
        }
        auto valid_bc = [](int b) { return b == MHD_HIP_PERIODIC || b == MHD_HIP_OUTFLOW; };
        if (!valid_bc(d.bc_x) || !valid_bc(d.bc_y)) {
            throw InvalidArgument("instance boundaries must be periodic or outflow");
        }
        if (scheme.limiter != MHD_HIP_MINMOD && scheme.limiter != MHD_HIP_MC) {
            throw InvalidArgument("unknown slope limiter");
        }
        // the stage kernels address a lane's cell by a 32-bit byte offset within its plane
        if (static_cast<unsigned long long>(d.nx + 4) * static_cast<unsigned long long>(d.ny + 4) *
                elem >=
            (1ull << 32)) {
            throw InvalidArgument("instance grid too large for the batch engine (a component "
                                  "plane of 4 GiB or more); use the single-run solver");
        }
        const size_t cells = static_cast<size_t>(d.nx) * d.ny;
        if (mask != nullptr && mask_len != cells) {
            throw InvalidArgument("FluidMask size does not match Grid2D interior");
        }
        // solver_api.hpp:300: an all-fluid mask is no mask
        bool masked = false;
        for (size_t k = 0; mask != nullptr && k < cells && !masked; ++k) {
            masked = mask[k] == 0;
        }
        std::vector<long long> plan;
        if (masked && walls) {
            const mhd_hip_grid grid{d.nx, d.ny, 2, d.x_min, d.x_max, d.y_min, d.y_max};
            mhdhip::wall_plan(grid, mask, plan);
        }
        const size_t plan_count = plan.size() / 3;
        int si = -1;
        for (int k = 0; k < max_instances; ++k) {
            if (!slots[k].used) {
                si = k;
                break;
            }
        }
        if (si < 0) {
            throw Full("all instance slots are in use");
        }
        const size_t bytes = instance_bytes(precision, d.nx, d.ny, masked, plan_count);
        size_t off = 0;
        if (!arena.alloc(bytes, off)) {
            throw Full("HBM budget exhausted (" + std::to_string(arena.used()) + " of " +
                       std::to_string(arena.capacity()) + " bytes in use)");
        }
        check(hipSetDevice(device), "hipSetDevice");
        Slot& s = slots[si];
        s = Slot{};
        s.used = true;
        s.id = next_id++;
        s.off = off;
        s.bytes = bytes;
        s.state_bytes = state_bytes_of(precision, d.nx, d.ny);
        s.desc = d;
        s.limiter = scheme.limiter;
        s.ni = nonideal_of(scheme);
        s.masked = masked;
        s.walls = walls;
        s.plan_count = static_cast<int>(plan_count);
        try {
            if (masked) {
                // mask and plan from host memory that lives until start_instance synchronises
                check(hipMemcpyAsync(const_cast<uint8_t*>(mask_of(s)), mask, cells,
                                     hipMemcpyHostToDevice, stream),
                      "hipMemcpyAsync(mask)");
                if (plan_count > 0) {
                    check(hipMemcpyAsync(const_cast<long long*>(plan_of(s)), plan.data(),
                                         plan.size() * sizeof(long long), hipMemcpyHostToDevice,
                                         stream),
                          "hipMemcpyAsync(plan)");
                }
            }
            if (precision == MHD_HIP_FP64) {
                upload_state<double>(s, comps);
                start_instance<double, 0>(si);
            } else if (precision == MHD_HIP_FP64_FAST) {
                upload_state<double>(s, comps);
                start_instance<double, 1>(si);
            } else {
                upload_state<float>(s, comps);
                start_instance<float, 0>(si);
            }
        } catch (...) {
            (void)hipStreamSynchronize(stream);
            arena.release(off, bytes);
            s = Slot{};
            throw;
        }
        dirty = true;
        return s.id;
    }

    void remove(int id)
    {
        Slot& s = slot_of(id);
        check(hipSetDevice(device), "hipSetDevice");
        // in-flight replays may still read this slot's state: release only after them
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        arena.release(s.off, s.bytes);
        s = Slot{};
        dirty = true;
    }

    // Row lengths the load model weighs (and the timed choice shortlists from).
    static constexpr int kModelRowCandidates[] = {8, 12, 16, 20, 24, 28, 32, 40, 48, 64, 96, 128};

    // waves of variant g's persistent grid in stage st (1 or 2)
    int group_waves(int g, int st) const
    {
        return group_blocks[g][st - 1] * mhdhip::kWavesPerBlock;
    }

    // The stepping instances of kernel variant g.
    std::vector<mhdhip::PlanInst> stepping_set(int g) const
    {
        std::vector<mhdhip::PlanInst> insts;
        for (int si = 0; si < max_instances; ++si) {
            const Slot& s = slots[si];
            if (s.used && !s.finished && !s.paused && group_of(s) == g) {
                insts.push_back(mhdhip::PlanInst{si, s.desc.nx, s.desc.ny});
            }
        }
        return insts;
    }

    // The partition a mode stands for on a stepping set: > 0 fixed segments of that many rows
    // in per-XCD queues (0: the load model's length for the set); < 0 the capacity fill with
    // pieces of at least -mode rows, as static per-wave lists.
    static int resolve(int mode, const std::vector<mhdhip::PlanInst>& insts, int waves)
    {
        if (mode != 0) return mode;
        std::vector<int> cands;
        for (const int r : kModelRowCandidates) {
            if (r >= mhdhip::kPickMinRows) cands.push_back(r);
        }
        return mhdhip::model_rows_for(insts, waves, cands, mhdhip::HandOut::kQueued,
                                      mhdhip::kPickTileOverhead);
    }

    // Before a sequence of stage launches: the count of tabled instances still to step (from
    // the control blocks as last read; the stream has drained since, so the pinned word is
    // free), and the queue counters zeroed if an earlier sequence was cut short
    // (counters_known is set again once the sequence is enqueued whole).
    void open_launches()
    {
        int live = 0;
        for (int si = 0; si < max_instances; ++si) {
            const Slot& s = slots[si];
            live += (tabled[si] && s.used && h_ctrl[si].t[0] < s.desc.t_end &&
                     h_ctrl[si].steps[0] < s.desc.max_steps)
                        ? 1
                        : 0;
        }
        h_live[0] = h_live[1] = live;  // every sequence starts at parity 0
        check(hipMemcpyAsync(d_live, h_live, 2 * sizeof(int), hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync(live)");
        if (!counters_known) {
            check(hipMemsetAsync(d_next, 0,
                                 sizeof(int) * static_cast<size_t>(kGroups) * 2 *
                                     mhdhip::kPlanXcds,
                                 stream),
                  "hipMemsetAsync(next)");
        }
        counters_known = false;
    }

    int plan_mode() const
    {
        if (seg_rows > 0 || seg_rows < -1) return seg_rows;
        if (seg_rows == -1 && tuned_rows != 0) return tuned_rows;
        return 0;
    }

    // Tables of the stepping set: every slot's BatchInst; per kernel variant and stage, its
    // tiles, as per-XCD queues (offsets at (2 g + st - 1) kQueueSlots) for fixed row lengths or
    // as per-wave lists (offsets at (2 g + st - 1) (max_waves + 1)) for the capacity fill; the
    // list of instances whose ghosts and solid cells are filled before each stage. The two
    // stages of a variant share one tile list when their grids split it alike.
    void build_tables(int mode)
    {
        std::vector<BatchTile> tiles;
        std::vector<int> first(static_cast<size_t>(kGroups) * 2 * (max_waves + 1), 0);
        std::vector<int> qbegin(static_cast<size_t>(kGroups) * 2 * kQueueSlots, 0);
        used_groups = 0;
        last_rows = 0;
        queued = false;
        std::fill(tabled.begin(), tabled.end(), 0);
        for (int g = 0; g < kGroups; ++g) {
            const std::vector<mhdhip::PlanInst> insts = stepping_set(g);
            for (const mhdhip::PlanInst& in : insts) tabled[static_cast<size_t>(in.slot)] = 1;
            for (int st = 1; st <= 2; ++st) {
                const int base = static_cast<int>(tiles.size());
                int* f = first.data() + static_cast<size_t>(2 * g + st - 1) * (max_waves + 1);
                int* qb = qbegin.data() + static_cast<size_t>(2 * g + st - 1) * kQueueSlots;
                std::fill(f, f + max_waves + 1, base);
                std::fill(qb, qb + kQueueSlots, base);
                if (insts.empty()) continue;
                const int waves = group_waves(g, st);
                const int r = resolve(mode, insts, waves);
                if (r > 0) {
                    const int nq = mhdhip::launch_queues(group_blocks[g][st - 1]);
                    if (st == 2 && nq == mhdhip::launch_queues(group_blocks[g][0])) {
                        std::copy(qb - kQueueSlots, qb, qb);  // stage 1's queues
                    } else {
                        const mhdhip::QueuePlan p = mhdhip::plan_queues(insts, r, nq);
                        for (int x = 0; x <= nq; ++x) qb[x] = base + p.begin[static_cast<size_t>(x)];
                        tiles.insert(tiles.end(), p.tiles.begin(), p.tiles.end());
                    }
                    queued = true;
                } else {
                    if (st == 2 && waves == group_waves(g, 1)) {
                        std::copy(f - (max_waves + 1), f, f);  // stage 1's list
                        continue;
                    }
                    const mhdhip::WavePlan p = mhdhip::plan_fill(insts, waves, -r);
                    for (int w = 0; w <= waves; ++w) f[w] = base + p.first[static_cast<size_t>(w)];
                    tiles.insert(tiles.end(), p.tiles.begin(), p.tiles.end());
                }
                if (last_rows == 0 || st == 2) last_rows = r > 0 ? r : 0;
            }
            if (!insts.empty()) used_groups |= 1u << g;
        }
        int nfill = 0;
        const size_t n = tiles.size();
        ensure_tiles(std::max<size_t>(n, 1));
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");  // staging reuse
        for (int si = 0; si < max_instances; ++si) {
            const Slot& s = slots[si];
            BatchInst& in = h_insts[si];
            std::memset(&in, 0, sizeof(in));
            if (!s.used) {
                continue;
            }
            in.U = base_of(s);
            in.Us = base_of(s) + s.state_bytes;
            in.mask = mask_of(s);
            in.plan = s.walls ? plan_of(s) : nullptr;
            in.plan_count = s.walls ? s.plan_count : 0;
            in.g = geom_of(s.desc);
            in.ph = physics_of(s);
            in.enabled = 1;
            in.stepping = tabled[static_cast<size_t>(si)];
            if (s.masked && !s.finished && !s.paused) {
                h_fill[1 + nfill++] = si;
            }
        }
        h_fill[0] = nfill;
        fill_needed = nfill > 0;
        std::copy(tiles.begin(), tiles.end(), h_tiles);
        std::copy(first.begin(), first.end(), h_first);
        std::copy(qbegin.begin(), qbegin.end(), h_queue);
        ntiles = static_cast<int>(n);
        check(hipMemcpyAsync(d_insts, h_insts, sizeof(BatchInst) * max_instances,
                             hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync(insts)");
        if (n > 0) {
            check(hipMemcpyAsync(d_tiles, h_tiles, sizeof(BatchTile) * n, hipMemcpyHostToDevice,
                                 stream),
                  "hipMemcpyAsync(tiles)");
        }
        check(hipMemcpyAsync(d_first, h_first, sizeof(int) * first.size(), hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync(first)");
        check(hipMemcpyAsync(d_queue, h_queue, sizeof(int) * qbegin.size(), hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync(queue)");
        check(hipMemcpyAsync(d_fill, h_fill, sizeof(int) * (nfill + 1), hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync(fill)");
    }

    BatchTables tables(int g, int st) const
    {
        const size_t k = static_cast<size_t>(2 * g + st - 1);
        const size_t other = static_cast<size_t>(2 * g + (st == 1 ? 1 : 0));
        return BatchTables{d_insts,
                           d_ctrl,
                           d_tiles,
                           d_first + k * (max_waves + 1),
                           d_fill,
                           d_fill + 1,
                           dry_run,
                           queued ? d_queue + k * kQueueSlots : nullptr,
                           d_next + k * mhdhip::kPlanXcds,
                           d_next + other * mhdhip::kPlanXcds,
                           d_live,
                           max_instances};
    }

    // One RK stage of the stepping set: the ghost and wall fill of the masked instances
    // (include/mhd/time_integrator.hpp:46-56), then one persistent launch per kernel variant.
    template <typename R, int FAST>
    void launch_stage(int stage, int parity)
    {
        if (fill_needed) {
            check(mhdhip::launch_batch_fill<R>(tables(0, stage), parity, stage, max_instances,
                                               stream),
                  "batch fill");
        }
        for (int g = 0; g < kGroups; ++g) {
            if (used_groups & (1u << g)) {
                check(mhdhip::launch_batch_stage<R, FAST>(stage, group_limiter(g), group_nonideal(g),
                                                    group_remap(g), tables(g, stage), parity,
                                                    group_blocks[g][stage - 1], stream),
                      stage == 1 ? "batch stage1" : "batch stage2");
            }
        }
    }

    unsigned launch_key() const
    {
        return used_groups | (fill_needed ? 1u << kGroups : 0u) |
               (queued ? 1u << (kGroups + 1) : 0u);
    }

    // The captured graph of kSteps steps; captured again when the kernel variants in use, the
    // need of the fill or the kind of hand-out change, since the graph holds one launch per
    // variant with the hand-out's table pointers.
    template <typename R, int FAST>
    void ensure_graph()
    {
        if (graph && graph_key == launch_key()) {
            return;
        }
        if (graph) {
            (void)hipGraphExecDestroy(graph);
            graph = nullptr;
        }
        hipGraph_t g = nullptr;
        check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "BeginCapture");
        try {
            for (int k = 0; k < kSteps; ++k) {
                const int p = k & 1;
                launch_stage<R, FAST>(1, p);
                launch_stage<R, FAST>(2, p);
            }
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check(hipStreamEndCapture(stream, &g), "EndCapture");
        const hipError_t e = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        check(e, "hipGraphInstantiate");
        graph_key = launch_key();
    }

    // Timed choice (rows per wave -1), made before the first replay of an engine whose choice
    // is not made yet: each shortlisted length is timed on dry steps, stage 1 as in a run (it
    // writes only U*) and stage 2 writing U* in place of U with no control field touched, so no
    // instance advances and the samples cost nothing but time. Shortlist: up to three fixed row
    // lengths, those of least load-model cost (summed over the kernel variants in use) within
    // 12 % of the least. Two sweeps over the shortlist, the second in reverse order, each
    // candidate one untimed step and three timed; the fastest step decides.
    std::vector<int> shortlist() const
    {
        std::vector<std::vector<mhdhip::PlanInst>> sets;
        for (int g = 0; g < kGroups; ++g) sets.push_back(stepping_set(g));
        std::vector<mhdhip::PlanLaunch> launches;
        for (int g = 0; g < kGroups; ++g) {
            for (int st = 1; st <= 2; ++st) {
                launches.push_back(mhdhip::PlanLaunch{&sets[g], group_waves(g, st)});
            }
        }
        return mhdhip::shortlist_rows_for(launches,
                                          std::vector<int>(std::begin(kModelRowCandidates),
                                                           std::end(kModelRowCandidates)),
                                          mhdhip::HandOut::kQueued);
    }

    template <typename R, int FAST>
    void autotune()
    {
        const std::vector<int> modes = shortlist();
        const int n = static_cast<int>(modes.size());
        std::vector<float> best(modes.size(), 1e30f);
        while (tev.size() < 2) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            tev.push_back(e);
        }
        dry_run = 1;
        try {
            for (int sweep = 0; sweep < 2; ++sweep) {
                for (int j = 0; j < n; ++j) {
                    const int i = sweep ? n - 1 - j : j;
                    build_tables(modes[i]);
                    open_launches();
                    launch_stage<R, FAST>(1, 0);
                    launch_stage<R, FAST>(2, 0);
                    counters_known = true;
                    for (int rep = 0; rep < 3; ++rep) {
                        check(hipEventRecord(tev[0], stream), "hipEventRecord");
                        open_launches();
                        launch_stage<R, FAST>(1, 0);
                        launch_stage<R, FAST>(2, 0);
                        counters_known = true;
                        check(hipEventRecord(tev[1], stream), "hipEventRecord");
                        check(hipEventSynchronize(tev[1]), "hipEventSynchronize");
                        float ms = 0.0f;
                        check(hipEventElapsedTime(&ms, tev[0], tev[1]), "hipEventElapsedTime");
                        best[i] = std::min(best[i], ms);
                    }
                }
            }
        } catch (...) {
            dry_run = 0;
            throw;
        }
        dry_run = 0;
        tuned_rows = modes[std::min_element(best.begin(), best.end()) - best.begin()];
        dirty = true;
    }

    void read_ctrl()
    {
        check(hipMemcpyAsync(h_ctrl, d_ctrl, sizeof(Ctrl) * max_instances, hipMemcpyDeviceToHost,
                             stream),
              "hipMemcpyAsync(ctrl)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (int si = 0; si < max_instances; ++si) {
            Slot& s = slots[si];
            if (!s.used || s.finished) {
                continue;
            }
            const Ctrl& c = h_ctrl[si];
            // replays hold an even number of steps: the current slot is parity 0
            s.result.t = c.t[0];
            s.result.steps = c.steps[0];
            s.result.c_h = c.ch_last[0];
            s.bad_dt = c.bad_dt;
            if (c.bad_dt || !(c.t[0] < s.desc.t_end && c.steps[0] < s.desc.max_steps)) {
                s.finished = true;
                dirty = true;  // drop its tiles from the next replay
            }
        }
    }

    // dt of the next step as the host can estimate it: the mean dt so far, or before the
    // first step the admission CFL reduction's (include/mhd/cfl.hpp:77-89)
    double dt_estimate(const Slot& s, const Ctrl& c) const
    {
        if (c.steps[0] > 0) {
            return c.t[0] / c.steps[0];
        }
        const Physics ph = physics_of(s);
        double ch;
        const unsigned long long bits = c.chmax[0];
        std::memcpy(&ch, &bits, sizeof(ch));
        double dt = ch > 0.0 ? ph.cfl / (ch * ph.inv_dx + ch * ph.inv_dy)
                             : std::numeric_limits<double>::infinity();
        if (ph.has_diff) dt = std::min(dt, ph.dt_diff);
        return dt;
    }

    void schedule()
    {
        if (sched_policy != MHD_HIP_SCHED_SRPT) {
            return;
        }
        struct Work {
            double w;
            int si;
            long long cells;
        };
        std::vector<Work> order;
        order.reserve(slots.size());
        for (int si = 0; si < max_instances; ++si) {
            const Slot& s = slots[si];
            if (!s.used || s.finished) {
                continue;
            }
            const Ctrl& c = h_ctrl[si];
            const double dt = dt_estimate(s, c);
            const double left = std::max(s.desc.t_end - c.t[0], 0.0);
            double steps = (dt > 0.0 && std::isfinite(dt)) ? left / dt : 0.0;
            steps = std::min(steps, static_cast<double>(s.desc.max_steps - c.steps[0]));
            const long long cells = static_cast<long long>(s.desc.nx) * s.desc.ny;
            order.push_back(Work{static_cast<double>(cells) * steps, si, cells});
        }
        std::stable_sort(order.begin(), order.end(),
                         [](const Work& a, const Work& b) { return a.w < b.w; });
        long long stepping = 0;
        for (const Work& o : order) {
            const bool run = stepping < sched_cap;
            stepping += run ? o.cells : 0;
            Slot& s = slots[o.si];
            if (s.paused == run) {
                s.paused = !run;
                dirty = true;
            }
        }
    }

    // Step counts and the stepping set of the replay just launched, for the progress check.
    struct Launched {
        std::vector<int> steps;
        std::vector<char> stepping;
        bool valid = false;
    };

    void note_launch(Launched& l) const
    {
        l.steps.assign(static_cast<size_t>(max_instances), 0);
        l.stepping.assign(static_cast<size_t>(max_instances), 0);
        for (int si = 0; si < max_instances; ++si) {
            const Slot& s = slots[si];
            l.stepping[si] = s.used && !s.finished && !s.paused;
            l.steps[si] = h_ctrl[si].steps[0];
        }
        l.valid = true;
    }

    // After read_ctrl: a replay that advanced none of the instances it stepped (and finished
    // none) would repeat forever; stop with an error instead.
    void check_progress(const Launched& l) const
    {
        if (!l.valid) return;
        std::vector<int> after(static_cast<size_t>(max_instances), 0);
        std::vector<char> fin(static_cast<size_t>(max_instances), 0);
        for (int si = 0; si < max_instances; ++si) {
            after[si] = h_ctrl[si].steps[0];
            fin[si] = slots[si].used && slots[si].finished;
        }
        if (!mhdhip::replay_progressed(l.steps, after, l.stepping, fin)) {
            throw std::runtime_error(
                "batch advance: a replay advanced none of the instances it stepped");
        }
    }

    template <typename R, int FAST>
    int advance_t(int max_replays)
    {
        check(hipSetDevice(device), "hipSetDevice");
        check(hipEventRecord(ev0, stream), "hipEventRecord");
        int active = 0;
        Launched last;
        for (int rep = 0; max_replays < 0 || rep < max_replays; ++rep) {
            read_ctrl();
            check_progress(last);
            schedule();
            active = 0;
            for (const auto& s : slots) {
                active += (s.used && !s.finished && !s.paused) ? 1 : 0;
            }
            if (active == 0) {
                last.valid = false;
                break;
            }
            if (seg_rows == -1 && tuned_rows == 0) {
                autotune<R, FAST>();
            }
            note_launch(last);
            if (dirty) {
                build_tables(plan_mode());
                dirty = false;
            }
            if (kernel_timing) {
                timed_replay<R, FAST>();
            } else {
                ensure_graph<R, FAST>();
                open_launches();
                check(hipGraphLaunch(graph, stream), "hipGraphLaunch");
                counters_known = true;
            }
        }
        check(hipEventRecord(ev1, stream), "hipEventRecord");
        check(hipEventSynchronize(ev1), "hipEventSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
        last_seconds = ms * 1e-3;
        read_ctrl();
        check_progress(last);
        active = 0;
        for (const auto& s : slots) {
            active += (s.used && !s.finished && !s.paused) ? 1 : 0;
        }
        return active;
    }

    // What the next replay needs before its launch, done now: the control blocks read, the
    // stepping set scheduled, the timed row choice made, the tile tables built and the step
    // graph captured. The next advance then starts with the launch.
    template <typename R, int FAST>
    void prepare_t()
    {
        check(hipSetDevice(device), "hipSetDevice");
        read_ctrl();
        schedule();
        bool any = false;
        for (const auto& s : slots) {
            any = any || (s.used && !s.finished && !s.paused);
        }
        if (!any) {
            return;
        }
        if (seg_rows == -1 && tuned_rows == 0) {
            autotune<R, FAST>();
        }
        if (dirty) {
            build_tables(plan_mode());
            dirty = false;
        }
        if (!kernel_timing) {
            ensure_graph<R, FAST>();
        }
    }

    void prepare()
    {
        switch (precision) {
            case MHD_HIP_FP64: prepare_t<double, 0>(); break;
            case MHD_HIP_FP64_FAST: prepare_t<double, 1>(); break;
            default: prepare_t<float, 0>(); break;
        }
    }

    int advance(int max_replays)
    {
        switch (precision) {
            case MHD_HIP_FP64: return advance_t<double, 0>(max_replays);
            case MHD_HIP_FP64_FAST: return advance_t<double, 1>(max_replays);
            default: return advance_t<float, 0>(max_replays);
        }
    }

    // include/mhd/time_integrator.hpp:101-104: the final ghost fill of a solve, and the wall
    // fill of its solid cells
    template <typename R>
    void final_fill(const Slot& s)
    {
        const Geom g = geom_of(s.desc);
        R* u = reinterpret_cast<R*>(base_of(s));
        const Gate open{d_ctrl, 0.0, 0, -1};
        check(mhdhip::launch_ghost_fill<R>(u, g, open, stream), "ghost_fill");
        if (s.masked && s.walls) {
            check(mhdhip::launch_wall_fill<R>(u, g, plan_of(s), s.plan_count, s.desc.gamma, open,
                                              stream),
                  "wall_fill");
        }
    }

    template <typename R>
    void download_t(const Slot& s, double* const comps[9])
    {
        R* u = reinterpret_cast<R*>(base_of(s));
        final_fill<R>(s);
        const size_t n = state_elems(s.desc.nx, s.desc.ny);
        ensure_stage(9 * n * sizeof(R));
        check(hipMemcpyAsync(h_stage, u, 9 * n * sizeof(R), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(download)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        mhdhip::scatter_planes(comps, reinterpret_cast<const R*>(h_stage), 9, n);
    }

    // Relative L1 distance per component of instance `id` from instance `ref_id` of `ref`
    // (the reference; possibly this engine), over the interior, on this engine's stream after
    // `ref`'s pending work.
    void rel_l1(int id, mhd_hip_batch& ref, int ref_id, double out[9])
    {
        const Slot& s = slot_of(id);
        const Slot& r = ref.slot_of(ref_id);
        if (s.desc.nx != r.desc.nx || s.desc.ny != r.desc.ny) {
            throw InvalidArgument("rel_l1: the two instances must have the same grid");
        }
        if (device != ref.device) {
            throw InvalidArgument("rel_l1: the two engines are on different devices");
        }
        check(hipSetDevice(device), "hipSetDevice");
        if (!l1_partial) {
            check(hipMalloc(reinterpret_cast<void**>(&l1_partial),
                            sizeof(mhdhip::DD) * 18 * mhdhip::kTotalsBlocks),
                  "hipMalloc(rel_l1)");
        }
        if (!totals_out) {
            check(hipMalloc(reinterpret_cast<void**>(&totals_out), 9 * sizeof(double)),
                  "hipMalloc(totals)");
        }
        check(hipStreamSynchronize(ref.stream), "hipStreamSynchronize");
        const void* a = static_cast<const char*>(arena_base) + s.off;
        const void* b = static_cast<const char*>(ref.arena_base) + r.off;
        check(mhdhip::launch_rel_l1_any(a, precision != MHD_HIP_FP32, b,
                                        ref.precision != MHD_HIP_FP32, geom_of(s.desc),
                                        l1_partial, totals_out, stream),
              "rel_l1");
        check(hipMemcpyAsync(out, totals_out, 9 * sizeof(double), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(rel_l1)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void totals(int id, double out[9])
    {
        const Slot& s = slot_of(id);
        check(hipSetDevice(device), "hipSetDevice");
        if (!totals_partial) {
            check(hipMalloc(reinterpret_cast<void**>(&totals_partial),
                            sizeof(mhdhip::DD) * 9 * mhdhip::kTotalsBlocks),
                  "hipMalloc(totals)");
            check(hipMalloc(reinterpret_cast<void**>(&totals_out), 9 * sizeof(double)),
                  "hipMalloc(totals)");
        }
        const Geom g = geom_of(s.desc);
        void* u = static_cast<char*>(arena_base) + s.off;
        if (precision != MHD_HIP_FP32) {
            check(mhdhip::launch_totals<double>(static_cast<double*>(u), g, totals_partial,
                                                totals_out, stream),
                  "totals");
        } else {
            check(mhdhip::launch_totals<float>(static_cast<float*>(u), g, totals_partial,
                                               totals_out, stream),
                  "totals");
        }
        check(hipMemcpyAsync(out, totals_out, 9 * sizeof(double), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(totals)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // max |div B| of one instance after its final ghost fill (as download sees the state)
    template <typename R>
    double max_div_b_t(const Slot& s)
    {
        const Geom g = geom_of(s.desc);
        const Physics ph = physics_of(s);
        R* u = reinterpret_cast<R*>(base_of(s));
        final_fill<R>(s);
        auto* out = reinterpret_cast<unsigned long long*>(totals_out);
        check(mhdhip::launch_max_div_b<R>(u, g, ph.inv_2dx, ph.inv_2dy, out, stream), "max_div_b");
        double v = 0.0;
        check(hipMemcpyAsync(&v, totals_out, sizeof(double), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(max_div_b)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        return v;
    }

    double max_div_b(int id)
    {
        const Slot& s = slot_of(id);
        check(hipSetDevice(device), "hipSetDevice");
        if (!totals_out) {
            check(hipMalloc(reinterpret_cast<void**>(&totals_out), 9 * sizeof(double)),
                  "hipMalloc(totals)");
        }
        return precision != MHD_HIP_FP32 ? max_div_b_t<double>(s) : max_div_b_t<float>(s);
    }

    void download(int id, double* const comps[9], size_t n)
    {
        const Slot& s = slot_of(id);
        if (n != state_elems(s.desc.nx, s.desc.ny)) {
            throw InvalidArgument("download: size does not match the instance's full layout");
        }
        check(hipSetDevice(device), "hipSetDevice");
        if (precision != MHD_HIP_FP32) {
            download_t<double>(s, comps);
        } else {
            download_t<float>(s, comps);
        }
    }
};

extern "C" {

size_t mhd_hip_batch_instance_bytes(int precision, int nx, int ny)
{
    if (nx < 1 || ny < 1 ||
        (precision != MHD_HIP_FP64 && precision != MHD_HIP_FP32 && precision != MHD_HIP_FP64_FAST)) {
        return 0;
    }
    return mhd_hip_batch::instance_bytes(precision, nx, ny);
}

int mhd_hip_batch_create(int precision, const mhd_hip_params* physics, size_t hbm_budget_bytes,
                         int max_instances, mhd_hip_batch** out)
{
    return guarded([&] {
        if (!physics || !out) throw InvalidArgument("null argument");
        *out = nullptr;
        auto b = std::make_unique<mhd_hip_batch>();
        b->init(precision, *physics, hbm_budget_bytes, max_instances);
        *out = b.release();
    });
}

void mhd_hip_batch_destroy(mhd_hip_batch* b)
{
    delete b;
}

int mhd_hip_batch_add(mhd_hip_batch* b, const mhd_hip_instance* inst,
                      const double* const comps[9], int* id)
{
    return guarded([&] {
        if (!b || !inst || !comps || !id) throw InvalidArgument("null argument");
        for (int c = 0; c < 9; ++c) {
            if (!comps[c]) throw InvalidArgument("null component pointer");
        }
        *id = b->add(*inst, b->phys_template, nullptr, 0, true, comps);
    });
}

int mhd_hip_batch_add_config(mhd_hip_batch* b, const mhd_hip_instance* inst,
                             const mhd_hip_params* scheme, const uint8_t* mask, size_t mask_len,
                             int reflective_walls, const double* const comps[9], int* id)
{
    return guarded([&] {
        if (!b || !inst || !scheme || !comps || !id) throw InvalidArgument("null argument");
        for (int c = 0; c < 9; ++c) {
            if (!comps[c]) throw InvalidArgument("null component pointer");
        }
        *id = b->add(*inst, *scheme, mask, mask_len, reflective_walls != 0, comps);
    });
}

int mhd_hip_batch_remove(mhd_hip_batch* b, int id)
{
    return guarded([&] {
        if (!b) throw InvalidArgument("null batch");
        b->remove(id);
    });
}

int mhd_hip_batch_advance(mhd_hip_batch* b, int max_replays, int* active)
{
    return guarded([&] {
        if (!b || !active) throw InvalidArgument("null argument");
        *active = b->advance(max_replays);
    });
}

int mhd_hip_batch_prepare(mhd_hip_batch* b)
{
    return guarded([&] {
        if (!b) throw InvalidArgument("null batch");
        b->prepare();
    });
}

int mhd_hip_batch_result(mhd_hip_batch* b, int id, mhd_hip_result* result, int* finished)
{
    return guarded([&] {
        if (!b || !result || !finished) throw InvalidArgument("null argument");
        const Slot& s = b->slot_of(id);
        *result = s.result;
        *finished = s.finished ? 1 : 0;
        if (s.bad_dt) {
            throw InvalidArgument("ssp_rk2_step requires positive dt");
        }
        if (s.finished && s.result.t < s.desc.t_end && s.result.steps >= s.desc.max_steps &&
            s.desc.t_end != std::numeric_limits<double>::infinity()) {
            throw std::runtime_error("solve: reached max_steps before t_end");
        }
    });
}

int mhd_hip_batch_download(mhd_hip_batch* b, int id, double* const comps[9], size_t n)
{
    return guarded([&] {
        if (!b || !comps) throw InvalidArgument("null argument");
        for (int c = 0; c < 9; ++c) {
            if (!comps[c]) throw InvalidArgument("null component pointer");
        }
        b->download(id, comps, n);
    });
}

int mhd_hip_batch_rel_l1(mhd_hip_batch* a, int id_a, mhd_hip_batch* ref, int id_ref,
                         double l1[9])
{
    return guarded([&] {
        if (!a || !ref || !l1) throw InvalidArgument("null argument");
        a->rel_l1(id_a, *ref, id_ref, l1);
    });
}

int mhd_hip_batch_totals(mhd_hip_batch* b, int id, double sums[9])
{
    return guarded([&] {
        if (!b || !sums) throw InvalidArgument("null argument");
        b->totals(id, sums);
    });
}

int mhd_hip_batch_set_seg_rows(mhd_hip_batch* b, int rows)
{
    return guarded([&] {
        if (!b) throw InvalidArgument("null batch");
        b->seg_rows = rows;
        b->tuned_rows = 0;
        b->dirty = true;
    });
}

int mhd_hip_batch_max_div_b(mhd_hip_batch* b, int id, double* value)
{
    return guarded([&] {
        if (!b || !value) throw InvalidArgument("null argument");
        *value = b->max_div_b(id);
    });
}

int mhd_hip_batch_set_paused(mhd_hip_batch* b, int id, int paused)
{
    return guarded([&] {
        if (!b) throw InvalidArgument("null batch");
        Slot& s = b->slot_of(id);
        if (s.paused != (paused != 0)) {
            s.paused = paused != 0;
            b->dirty = true;
        }
    });
}

int mhd_hip_batch_progress(mhd_hip_batch* b, int id, double* t, int* steps, double* dt_estimate)
{
    return guarded([&] {
        if (!b || !t || !steps || !dt_estimate) throw InvalidArgument("null argument");
        const Slot& s = b->slot_of(id);
        const Ctrl& c = b->h_ctrl[b->slot_index(s)];
        *t = c.t[0];
        *steps = c.steps[0];
        *dt_estimate = b->dt_estimate(s, c);
    });
}

int mhd_hip_batch_set_schedule(mhd_hip_batch* b, int policy, long long work_cap_cells)
{
    return guarded([&] {
        if (!b) throw InvalidArgument("null batch");
        if (policy == MHD_HIP_SCHED_ALL) {
            for (auto& s : b->slots) {
                if (s.paused) {
                    s.paused = false;
                    b->dirty = true;
                }
            }
        } else if (policy == MHD_HIP_SCHED_SRPT) {
            if (work_cap_cells < 1) throw InvalidArgument("work cap must be at least one cell");
        } else {
            throw InvalidArgument("unknown scheduling policy " + std::to_string(policy));
        }
        b->sched_policy = policy;
        b->sched_cap = work_cap_cells;
    });
}

int mhd_hip_batch_poll_finished(mhd_hip_batch* b, int* ids, int cap, int* count)
{
    return guarded([&] {
        if (!b || !count || (cap > 0 && !ids)) throw InvalidArgument("null argument");
        int n = 0;
        for (const auto& s : b->slots) {
            if (s.used && s.finished && n < cap) {
                ids[n++] = s.id;
            }
        }
        *count = n;
    });
}

int mhd_hip_batch_set_kernel_timing(mhd_hip_batch* b, int enable)
{
    return guarded([&] {
        if (!b) throw InvalidArgument("null batch");
        b->kernel_timing = enable != 0;
        b->stage_ms[0] = b->stage_ms[1] = 0.0;
        b->stage_launches[0] = b->stage_launches[1] = 0;
    });
}

int mhd_hip_batch_kernel_times(const mhd_hip_batch* b, double* total_ms, long long* launches)
{
    return guarded([&] {
        if (!b || !total_ms || !launches) throw InvalidArgument("null argument");
        for (int k = 0; k < 2; ++k) {
            total_ms[k] = b->stage_ms[k];
            launches[k] = b->stage_launches[k];
        }
    });
}

int mhd_hip_batch_info(const mhd_hip_batch* b, mhd_hip_batch_stats* st)
{
    return guarded([&] {
        if (!b || !st) throw InvalidArgument("null argument");
        int n = 0, act = 0;
        for (const auto& s : b->slots) {
            n += s.used ? 1 : 0;
            act += (s.used && !s.finished) ? 1 : 0;
        }
        int paused = 0;
        for (const auto& s : b->slots) {
            paused += (s.used && !s.finished && s.paused) ? 1 : 0;
        }
        st->instances = n;
        st->active = act - paused;
        st->paused = paused;
        st->tiles = b->ntiles;
        st->seg_rows = b->last_rows;
        st->blocks = b->blocks;
        st->bytes_used = b->arena.used();
        st->bytes_budget = b->arena.capacity();
        st->last_seconds = b->last_seconds;
    });
}

}  // extern "C"
