// Single-run device solver and the C ABI of include/mhd_hip.h.
//
// One mhd_hip_solver owns the device state of one grid: U, U* (9 planes each, fp64 or fp32),
// the control block, the interior mask and the wall-fill plan. A solve runs on the solver's own
// stream: the first CFL reduction, then batches of steps replayed from a captured hipGraph
// (kSteps steps per replay, parity baked in), checking the device control block once per
// replay while the next replay is already queued. Surplus steps after t_end are no-ops inside
// the kernels, so the host never needs the step count in advance.
#include "mhd_abi.hpp"
#include "mhd_hip.h"
#include "mhd_host.hpp"
#include "mhd_launch.hpp"
#include "mhd_transfer.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <stdexcept>
#include <string>
#include <vector>

namespace mhdhip {

std::string& last_error_slot()
{
    thread_local std::string msg;
    return msg;
}

}  // namespace mhdhip

namespace {

using mhdhip::check;
using mhdhip::guarded;
using mhdhip::InvalidArgument;
using mhdhip::NoDevice;

// Process-wide autotuning results: best rows-per-wave for a (shape, precision, kernel variant).
struct TuneKey {
    int nx, ny, precision, limiter, ni, remap;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(nx, ny, precision, limiter, ni, remap) <
               std::tie(o.nx, o.ny, o.precision, o.limiter, o.ni, o.remap);
    }
};
std::mutex g_tune_mutex;
std::map<TuneKey, int> g_tuned;

bool device_is_gfx950(int dev)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        return false;
    }
    return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

}  // namespace

namespace {

// Scratch of mhd_hip_totals_noalloc: the per-workgroup partial sums of k_totals_partial
// (kTotalsBlocks x 9 double-double values, written in full by every call) and the nine
// totals k_totals_final writes, 73,800 bytes per device for the life of the process. Host
// code never passes `scratch` itself to a kernel or a copy: it is a host-side stand-in;
// noalloc_scratch(device) gives that device's address of the array.
constexpr size_t kNoallocPartials = static_cast<size_t>(mhdhip::kTotalsBlocks) * mhdhip::kNC;
constexpr size_t kNoallocBytes = kNoallocPartials * sizeof(mhdhip::DD) + 9 * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mutex;
void* g_noalloc_ptr[kNoallocMaxDevices] = {};

// The device address of `scratch` on `device`, looked up once per device with that device
// current (each device holds its own copy of the array). Caller holds g_noalloc_mutex.
void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kNoallocMaxDevices) {
        throw InvalidArgument("totals without allocation: device index " + std::to_string(device) +
                              " beyond the " + std::to_string(kNoallocMaxDevices) +
                              " devices the scratch table holds");
    }
    if (g_noalloc_ptr[device] == nullptr) {
        check(hipSetDevice(device), "hipSetDevice");
        void* ptr = nullptr;
        check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (ptr == nullptr) {
            throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
        }
        g_noalloc_ptr[device] = ptr;
    }
    return g_noalloc_ptr[device];
}

}  // namespace

using mhdhip::Ctrl;
using mhdhip::Gate;
using mhdhip::Geom;
using mhdhip::Physics;
using mhdhip::StageBuffers;
using mhdhip::Tiling;

struct mhd_hip_solver {
    static constexpr int kSteps = 8;  // steps per captured graph (even: parity 0,1,...)

    int device = 0;
    mhd_hip_grid grid{};
    int precision = MHD_HIP_FP64;
    Geom geom{};
    double dx = 0.0, dy = 0.0;
    size_t elem = 8;

    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    void* U = nullptr;
    void* Us = nullptr;
    Ctrl* ctrl = nullptr;
    Ctrl* ctrl_host = nullptr;  // pinned
    double* dt_log = nullptr;
    int dt_cap = 0;
    int dt_count = 0;

    int bc_x = MHD_HIP_PERIODIC, bc_y = MHD_HIP_PERIODIC;
    bool masked = false;
    bool walls = true;
    uint8_t* mask_dev = nullptr;
    long long* plan_dev = nullptr;
    int plan_count = 0;

    int seg_rows = 0;   // rows marched per wave; 0 = autotuned per shape
    int last_rows = 0;  // value used by the last solve
    double last_seconds = 0.0;

    // kernel timing mode: fixed-step runs launch eagerly with events around each stage kernel
    bool kernel_timing = false;
    double stage_ms[2] = {0.0, 0.0};
    long long stage_launches[2] = {0, 0};
    std::vector<hipEvent_t> timing_events;

    hipGraphExec_t graph_exec = nullptr;
    std::vector<unsigned char> graph_key;

    mhdhip::DD* totals_partial = nullptr;  // compensated-sum scratch, allocated on first use
    double* totals_out = nullptr;

    mhdhip::StagedCopier staging;  // pinned chunk buffers for large uploads and downloads
    mhdhip::DD* l1_partial = nullptr;  // state-comparison scratch, allocated on first use

    ~mhd_hip_solver() { release(); }

    void release()
    {
        if (device >= 0) {
            (void)hipSetDevice(device);
        }
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (U) (void)hipFree(U);
        if (Us) (void)hipFree(Us);
        if (ctrl) (void)hipFree(ctrl);
        if (ctrl_host) (void)hipHostFree(ctrl_host);
        if (dt_log) (void)hipFree(dt_log);
        if (mask_dev) (void)hipFree(mask_dev);
        if (plan_dev) (void)hipFree(plan_dev);
        if (totals_partial) (void)hipFree(totals_partial);
        if (totals_out) (void)hipFree(totals_out);
        if (l1_partial) (void)hipFree(l1_partial);
        totals_partial = nullptr;
        totals_out = nullptr;
        l1_partial = nullptr;
        staging.release();
        for (hipEvent_t e : timing_events) (void)hipEventDestroy(e);
        timing_events.clear();
        if (ev_start) (void)hipEventDestroy(ev_start);
        if (ev_stop) (void)hipEventDestroy(ev_stop);
        if (stream) (void)hipStreamDestroy(stream);
        graph_exec = nullptr;
        U = Us = nullptr;
        ctrl = nullptr;
        ctrl_host = nullptr;
        dt_log = nullptr;
        mask_dev = nullptr;
        plan_dev = nullptr;
        ev_start = ev_stop = nullptr;
        stream = nullptr;
    }

    size_t plane() const { return static_cast<size_t>(geom.plane); }
    size_t state_bytes() const { return 9 * plane() * elem; }

    void init(const mhd_hip_grid& g, int prec)
    {
        if (g.ng < 2) {
            throw InvalidArgument("MHDSolver requires ng >= 2");
        }
        if (g.nx < 1 || g.ny < 1) {
            throw InvalidArgument("grid must have at least one interior cell per axis");
        }
        if (prec != MHD_HIP_FP64 && prec != MHD_HIP_FP32 && prec != MHD_HIP_FP64_FAST) {
            throw InvalidArgument(
                "precision must be MHD_HIP_FP64, MHD_HIP_FP32 or MHD_HIP_FP64_FAST");
        }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
            throw NoDevice("no HIP device available");
        }
        check(hipGetDevice(&device), "hipGetDevice");
        if (!device_is_gfx950(device)) {
            throw NoDevice("current HIP device is not gfx950 (MI355X)");
        }
        grid = g;
        precision = prec;
        elem = prec == MHD_HIP_FP32 ? sizeof(float) : sizeof(double);
        geom.nx = g.nx;
        geom.ny = g.ny;
        geom.ng = g.ng;
        geom.nxt = g.nx + 2 * g.ng;
        geom.nyt = g.ny + 2 * g.ng;
        geom.plane = static_cast<long long>(geom.nxt) * geom.nyt;
        geom.bc_x = MHD_HIP_PERIODIC;
        geom.bc_y = MHD_HIP_PERIODIC;
        // include/grid/grid.hpp:24-25
        dx = (g.x_max - g.x_min) / static_cast<double>(g.nx);
        dy = (g.y_max - g.y_min) / static_cast<double>(g.ny);

        check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        check(hipEventCreate(&ev_start), "hipEventCreate");
        check(hipEventCreate(&ev_stop), "hipEventCreate");
        check(hipMalloc(&U, state_bytes()), "hipMalloc(U)");
        check(hipMalloc(&Us, state_bytes()), "hipMalloc(U*)");
        check(hipMalloc(reinterpret_cast<void**>(&ctrl), sizeof(Ctrl)), "hipMalloc(ctrl)");
        check(hipHostMalloc(reinterpret_cast<void**>(&ctrl_host), sizeof(Ctrl), 0),
              "hipHostMalloc(ctrl)");
        // zero both states so that never-uploaded ghosts are defined
        check(hipMemsetAsync(U, 0, state_bytes(), stream), "hipMemset");
        check(hipMemsetAsync(Us, 0, state_bytes(), stream), "hipMemset");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void configure(int bx, int by, const uint8_t* mask, size_t mask_len, int reflective)
    {
        auto valid = [](int b) { return b == MHD_HIP_PERIODIC || b == MHD_HIP_OUTFLOW; };
        if (bx == MHD_HIP_REFLECTIVE || by == MHD_HIP_REFLECTIVE) {
            // include/boundary_conditions/boundary_conditions.hpp:52-55
            throw InvalidArgument("Reflective boundary conditions are not implemented yet");
        }
        if (!valid(bx) || !valid(by)) {
            throw InvalidArgument("unknown boundary condition type");
        }
        check(hipSetDevice(device), "hipSetDevice");
        bc_x = bx;
        bc_y = by;
        geom.bc_x = bx;
        geom.bc_y = by;
        walls = reflective != 0;
        masked = false;
        plan_count = 0;
        if (mask != nullptr) {
            const size_t need = static_cast<size_t>(grid.nx) * grid.ny;
            if (mask_len != need) {
                throw InvalidArgument("FluidMask size does not match Grid2D interior");
            }
            masked = std::any_of(mask, mask + need, [](uint8_t v) { return v == 0; });
        }
        invalidate_graph();
        if (!masked) {
            return;
        }
        const size_t need = static_cast<size_t>(grid.nx) * grid.ny;
        if (mask_dev == nullptr) {
            check(hipMalloc(reinterpret_cast<void**>(&mask_dev), need), "hipMalloc(mask)");
        }
        copy_sync(mask_dev, mask, need, hipMemcpyHostToDevice, "hipMemcpy(mask)");
        std::vector<long long> plan;
        plan_count = static_cast<int>(mhdhip::wall_plan(grid, mask, plan));
        if (plan_dev) {
            (void)hipFree(plan_dev);
            plan_dev = nullptr;
        }
        if (plan_count > 0) {
            check(hipMalloc(reinterpret_cast<void**>(&plan_dev), plan.size() * sizeof(long long)),
                  "hipMalloc(plan)");
            copy_sync(plan_dev, plan.data(), plan.size() * sizeof(long long),
                      hipMemcpyHostToDevice, "hipMemcpy(plan)");
        }
    }

    // Host <-> device copy on the solver's stream, complete on return. Not hipMemcpy: a copy on
    // the legacy stream fails while another thread captures a graph of its own solver.
    void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                   const char* what) const
    {
        check(hipMemcpyAsync(dst, src, bytes, kind, stream), what);
        check(hipStreamSynchronize(stream), what);
    }

    void upload(const double* const comps[9], size_t n)
    {
        if (n != plane()) {
            throw InvalidArgument("upload: size does not match the full SoA layout");
        }
        check(hipSetDevice(device), "hipSetDevice");
        if (9 * n * elem >= mhdhip::StagedCopier::kMinBytes) {
            if (precision != MHD_HIP_FP32) {
                staging.upload(static_cast<double*>(U), plane(), comps, 9, n, stream);
            } else {
                staging.upload(static_cast<float*>(U), plane(), comps, 9, n, stream);
            }
            return;
        }
        if (precision != MHD_HIP_FP32) {
            for (int c = 0; c < 9; ++c) {
                copy_sync(static_cast<double*>(U) + c * plane(), comps[c], n * sizeof(double),
                          hipMemcpyHostToDevice, "hipMemcpy(upload)");
            }
        } else {
            std::vector<float> stage(n);
            for (int c = 0; c < 9; ++c) {
                for (size_t k = 0; k < n; ++k) {
                    stage[k] = static_cast<float>(comps[c][k]);
                }
                copy_sync(static_cast<float*>(U) + c * plane(), stage.data(), n * sizeof(float),
                          hipMemcpyHostToDevice, "hipMemcpy(upload)");
            }
        }
    }

    void download(double* const comps[9], size_t n)
    {
        if (n != plane()) {
            throw InvalidArgument("download: size does not match the full SoA layout");
        }
        check(hipSetDevice(device), "hipSetDevice");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (9 * n * elem >= mhdhip::StagedCopier::kMinBytes) {
            if (precision != MHD_HIP_FP32) {
                staging.download(comps, static_cast<const double*>(U), plane(), 9, n, stream);
            } else {
                staging.download(comps, static_cast<const float*>(U), plane(), 9, n, stream);
            }
            return;
        }
        if (precision != MHD_HIP_FP32) {
            for (int c = 0; c < 9; ++c) {
                copy_sync(comps[c], static_cast<double*>(U) + c * plane(), n * sizeof(double),
                          hipMemcpyDeviceToHost, "hipMemcpy(download)");
            }
        } else {
            std::vector<float> stage(n);
            for (int c = 0; c < 9; ++c) {
                copy_sync(stage.data(), static_cast<float*>(U) + c * plane(), n * sizeof(float),
                          hipMemcpyDeviceToHost, "hipMemcpy(download)");
                for (size_t k = 0; k < n; ++k) {
                    comps[c][k] = static_cast<double>(stage[k]);
                }
            }
        }
    }

    Physics physics(const mhd_hip_params& p, double t_end, int max_steps) const
    {
        Physics ph{};
        ph.gamma = p.gamma;
        ph.gm1 = p.gamma - 1.0;
        ph.inv_gm1 = 1.0 / (p.gamma - 1.0);  // include/mhd/kernels/cpu_rows.hpp:104
        ph.cfl = p.cfl;
        ph.inv_dx = 1.0 / dx;
        ph.inv_dy = 1.0 / dy;
        ph.inv_2dx = 0.5 / dx;  // include/mhd/nonideal.hpp:33-34
        ph.inv_2dy = 0.5 / dy;
        ph.glm_alpha = p.glm_alpha;
        ph.h = (dy < dx) ? dy : dx;
        // include/physics_config/mhd_config.hpp:97-110
        double eta = 0.0;
        auto mx = [](double a, double b) { return (a < b) ? b : a; };
        if (p.ohmic) eta = mx(eta, std::fabs(p.eta_ohm));
        if (p.hall) eta = mx(eta, std::fabs(p.eta_hall));
        if (p.ambipolar) eta = mx(eta, std::fabs(p.eta_ambipolar));
        ph.has_diff = eta > 0.0 ? 1 : 0;
        // include/mhd/cfl.hpp:84-88
        ph.dt_diff = ph.has_diff ? p.cfl * ph.h * ph.h / (2.0 * 2.0 * eta) : 0.0;
        ph.t_end = t_end;
        ph.max_steps = max_steps;
        ph.ni.ohmic = p.ohmic;
        ph.ni.hall = p.hall;
        ph.ni.ambipolar = p.ambipolar;
        ph.ni.eta_ohm = p.eta_ohm;
        ph.ni.eta_hall = p.eta_hall;
        ph.ni.eta_ambipolar = p.eta_ambipolar;
        return ph;
    }

    static bool nonideal_any(const mhd_hip_params& p)
    {
        // include/physics_config/mhd_config.hpp:90-94
        return (p.ohmic && p.eta_ohm != 0.0) || (p.hall && p.eta_hall != 0.0) ||
               (p.ambipolar && p.eta_ambipolar != 0.0);
    }

    template <typename R, int FAST>
    void enqueue_step(const Physics& ph, const Tiling& tl, int limiter, bool ni, int parity)
    {
        const bool fill = masked && walls;
        const bool remap = !fill;
        const Gate gate{ctrl, ph.t_end, ph.max_steps, parity};
        R* u = static_cast<R*>(U);
        R* us = static_cast<R*>(Us);
        StageBuffers b1{u, u, us, masked ? mask_dev : nullptr, ctrl, dt_log, dt_cap};
        StageBuffers b2{us, u, u, masked ? mask_dev : nullptr, ctrl, dt_log, dt_cap};
        if (fill) {
            check(mhdhip::launch_ghost_fill<R>(u, geom, gate, stream), "ghost_fill");
            check(mhdhip::launch_wall_fill<R>(u, geom, plan_dev, plan_count, ph.gamma, gate, stream),
                  "wall_fill");
        }
        check(mhdhip::launch_stage<R, FAST>(1, limiter, ni, remap, b1, geom, ph, tl, parity, stream),
              "stage1");
        if (fill) {
            check(mhdhip::launch_ghost_fill<R>(us, geom, gate, stream), "ghost_fill");
            check(mhdhip::launch_wall_fill<R>(us, geom, plan_dev, plan_count, ph.gamma, gate,
                                              stream),
                  "wall_fill");
        }
        check(mhdhip::launch_stage<R, FAST>(2, limiter, ni, remap, b2, geom, ph, tl, parity, stream),
              "stage2");
    }

    // Eager launches of `steps` steps with an event pair around every stage kernel; adds the
    // kernel times to stage_ms. Returns the number of steps launched (even).
    template <typename R, int FAST>
    long long timed_steps(const Physics& ph, const Tiling& tl, int limiter, bool ni, int steps)
    {
        const long long n = steps + (steps & 1);
        const size_t need = static_cast<size_t>(4 * n);
        while (timing_events.size() < need) {
            hipEvent_t e;
            check(hipEventCreate(&e), "hipEventCreate");
            timing_events.push_back(e);
        }
        const bool fill = masked && walls;
        const bool remap = !fill;
        R* u = static_cast<R*>(U);
        R* us = static_cast<R*>(Us);
        StageBuffers b1{u, u, us, masked ? mask_dev : nullptr, ctrl, dt_log, dt_cap};
        StageBuffers b2{us, u, u, masked ? mask_dev : nullptr, ctrl, dt_log, dt_cap};
        for (long long k = 0; k < n; ++k) {
            const int parity = static_cast<int>(k & 1);
            const Gate gate{ctrl, ph.t_end, ph.max_steps, parity};
            hipEvent_t* e = &timing_events[4 * k];
            if (fill) {
                check(mhdhip::launch_ghost_fill<R>(u, geom, gate, stream), "ghost_fill");
                check(mhdhip::launch_wall_fill<R>(u, geom, plan_dev, plan_count, ph.gamma, gate,
                                                  stream),
                      "wall_fill");
            }
            check(hipEventRecord(e[0], stream), "hipEventRecord");
            check(mhdhip::launch_stage<R, FAST>(1, limiter, ni, remap, b1, geom, ph, tl, parity, stream),
                  "stage1");
            check(hipEventRecord(e[1], stream), "hipEventRecord");
            if (fill) {
                check(mhdhip::launch_ghost_fill<R>(us, geom, gate, stream), "ghost_fill");
                check(mhdhip::launch_wall_fill<R>(us, geom, plan_dev, plan_count, ph.gamma, gate,
                                                  stream),
                      "wall_fill");
            }
            check(hipEventRecord(e[2], stream), "hipEventRecord");
            check(mhdhip::launch_stage<R, FAST>(2, limiter, ni, remap, b2, geom, ph, tl, parity, stream),
                  "stage2");
            check(hipEventRecord(e[3], stream), "hipEventRecord");
        }
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        for (long long k = 0; k < steps; ++k) {
            float a = 0.0f, b = 0.0f;
            check(hipEventElapsedTime(&a, timing_events[4 * k], timing_events[4 * k + 1]),
                  "hipEventElapsedTime");
            check(hipEventElapsedTime(&b, timing_events[4 * k + 2], timing_events[4 * k + 3]),
                  "hipEventElapsedTime");
            stage_ms[0] += a;
            stage_ms[1] += b;
            stage_launches[0] += 1;
            stage_launches[1] += 1;
        }
        return n;
    }

    // Autotuner: times one step's pair of stage kernels for candidate row-segment lengths on
    // the current state (two sweeps over the candidates, the second in reverse order) and keeps
    // the fastest per shape and kernel variant. Stage 1 writes U*
    // from U; the timed stage 2 reads U* and U and writes U* again, so U is left untouched
    // (U* is scratch until the next stage 1; the control fields the two launches write are
    // rewritten by the step's own stages). Runs after the first CFL reduction of a solve, so
    // the kernels see a live step.
    template <typename R, int FAST>
    int tuned_rows(const Physics& ph, int limiter, bool ni)
    {
        const bool remap = !(masked && walls);
        const TuneKey key{grid.nx, grid.ny, precision, limiter, ni ? 1 : 0, remap ? 1 : 0};
        {
            std::lock_guard<std::mutex> lk(g_tune_mutex);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) {
                return it->second;
            }
        }
        if (const char* env = std::getenv("MHD_HIP_SEG_ROWS")) {
            const int v = std::atoi(env);
            if (v > 0) {
                return v;
            }
        }
        if (static_cast<long long>(grid.nx) * grid.ny < 4096) {
            return mhdhip::default_seg_rows(grid.nx, grid.ny);
        }
        R* u = static_cast<R*>(U);
        R* us = static_cast<R*>(Us);
        StageBuffers b1{u, u, us, masked ? mask_dev : nullptr, ctrl, nullptr, 0};
        StageBuffers b2{us, u, us, masked ? mask_dev : nullptr, ctrl, nullptr, 0};
        hipEvent_t e0, e1;
        check(hipEventCreate(&e0), "hipEventCreate");
        check(hipEventCreate(&e1), "hipEventCreate");
        // candidates: fixed lengths plus the lengths that make the wave count just fill
        // k whole rounds of the device's concurrent wave slots (no half-empty last round);
        // grids too small to fill the slots with 8-row segments also try 2 to 6 rows, since
        // there a step lasts as long as one wave's march
        std::vector<int> cands = {8, 16, 24, 32, 48, 64, 96, 128};
        {
            hipDeviceProp_t prop;
            check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
            const int per_cu = mhdhip::stage_blocks_per_cu<R, FAST>(limiter, ni, remap);
            const long long slots =
                static_cast<long long>(per_cu) * mhdhip::kWavesPerBlock * prop.multiProcessorCount;
            const long long strips = (grid.nx + mhdhip::kStripCells - 1) / mhdhip::kStripCells;
            for (int k = 1; k <= 6; ++k) {
                const long long segs = std::max<long long>(1, k * slots / strips);
                const int rows = static_cast<int>((grid.ny + segs - 1) / segs);
                if (rows >= 2 && rows <= 512) {
                    cands.push_back(rows);
                }
            }
            if (strips * ((grid.ny + 7) / 8) < slots) {
                for (const int rows : {2, 3, 4, 6}) {
                    cands.push_back(rows);
                }
            }
            std::sort(cands.begin(), cands.end());
            cands.erase(std::unique(cands.begin(), cands.end()), cands.end());
        }
        cands.erase(std::remove_if(cands.begin(), cands.end(),
                                   [&](int c) { return c > grid.ny && c != 8; }),
                    cands.end());
        std::vector<float> cand_best(cands.size(), 1e30f);
        try {
            // two sweeps, the second in reverse order, so that a clock ramp favours no end
            for (size_t k = 0; k < 2 * cands.size(); ++k) {
                const size_t ci = k < cands.size() ? k : 2 * cands.size() - 1 - k;
                const int cand = cands[ci];
                const Tiling tl = mhdhip::make_tiling(grid.nx, grid.ny, cand);
                // one untimed pair, then the best of three
                check(mhdhip::launch_stage<R, FAST>(1, limiter, ni, remap, b1, geom, ph, tl, 0, stream),
                      "tune");
                check(mhdhip::launch_stage<R, FAST>(2, limiter, ni, remap, b2, geom, ph, tl, 0, stream),
                      "tune");
                float cand_ms = 1e30f;
                for (int rep = 0; rep < 3; ++rep) {
                    check(hipEventRecord(e0, stream), "hipEventRecord");
                    check(mhdhip::launch_stage<R, FAST>(1, limiter, ni, remap, b1, geom, ph, tl, 0,
                                                  stream),
                          "tune");
                    check(mhdhip::launch_stage<R, FAST>(2, limiter, ni, remap, b2, geom, ph, tl, 0,
                                                  stream),
                          "tune");
                    check(hipEventRecord(e1, stream), "hipEventRecord");
                    check(hipEventSynchronize(e1), "hipEventSynchronize");
                    float ms = 0.0f;
                    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                    cand_ms = std::min(cand_ms, ms);
                }
                cand_best[ci] = std::min(cand_best[ci], cand_ms);
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const int best = cands.empty()
                             ? mhdhip::default_seg_rows(grid.nx, grid.ny)
                             : cands[std::min_element(cand_best.begin(), cand_best.end()) -
                                     cand_best.begin()];
        std::lock_guard<std::mutex> lk(g_tune_mutex);
        g_tuned[key] = best;
        return best;
    }

    void invalidate_graph()
    {
        if (graph_exec) {
            (void)hipGraphExecDestroy(graph_exec);
            graph_exec = nullptr;
        }
        graph_key.clear();
    }

    template <typename R, int FAST>
    void ensure_graph(const Physics& ph, const Tiling& tl, int limiter, bool ni)
    {
        std::vector<unsigned char> key(sizeof(Physics) + sizeof(Tiling) + 2 * sizeof(int) +
                                       sizeof(void*));
        unsigned char* k = key.data();
        std::memcpy(k, &ph, sizeof(Physics));
        k += sizeof(Physics);
        std::memcpy(k, &tl, sizeof(Tiling));
        k += sizeof(Tiling);
        const int flags[2] = {limiter, ni ? 1 : 0};
        std::memcpy(k, flags, sizeof(flags));
        k += sizeof(flags);
        std::memcpy(k, &dt_log, sizeof(void*));
        if (graph_exec && key == graph_key) {
            return;
        }
        invalidate_graph();
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "BeginCapture");
        try {
            for (int s = 0; s < kSteps; ++s) {
                enqueue_step<R, FAST>(ph, tl, limiter, ni, s & 1);
            }
        } catch (...) {
            (void)hipStreamEndCapture(stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        check(hipStreamEndCapture(stream, &graph), "EndCapture");
        const hipError_t e = hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e, "hipGraphInstantiate");
        graph_key = key;
    }

    template <typename R, int FAST>
    mhd_hip_result run(const mhd_hip_params& p, double t_end, int max_steps, bool allow_max_steps)
    {
        if (t_end < 0.0) {
            throw InvalidArgument("t_end must be non-negative");
        }
        if (p.cfl <= 0.0) {
            throw InvalidArgument("CFL number must be positive");
        }
        if (p.limiter != MHD_HIP_MINMOD && p.limiter != MHD_HIP_MC) {
            throw InvalidArgument("unknown slope limiter");
        }
        if (max_steps < 0) {
            throw InvalidArgument("max_steps must be non-negative");
        }
        check(hipSetDevice(device), "hipSetDevice");
        const Physics ph = physics(p, t_end, max_steps);
        const bool ni = nonideal_any(p);
        R* u = static_cast<R*>(U);

        check(hipEventRecord(ev_start, stream), "hipEventRecord");
        check(mhdhip::launch_ctrl_init<R>(ctrl, stream), "ctrl_init");
        check(mhdhip::launch_cfl<R, FAST>(u, geom, ph, masked ? mask_dev : nullptr, ctrl, stream), "cfl");
        const int rows = seg_rows > 0 ? seg_rows : tuned_rows<R, FAST>(ph, p.limiter, ni);
        const Tiling tl = mhdhip::make_tiling(grid.nx, grid.ny, rows);
        last_rows = rows;

        long long launched = 0;
        const bool any_step = t_end > 0.0 && max_steps > 0;
        if (any_step && kernel_timing && allow_max_steps) {
            launched = timed_steps<R, FAST>(ph, tl, p.limiter, ni, max_steps);
        } else if (any_step) {
            ensure_graph<R, FAST>(ph, tl, p.limiter, ni);
            // Two replays in flight: queue replay k+1, then inspect the control block after k.
            hipEvent_t done[2];
            check(hipEventCreateWithFlags(&done[0], hipEventDisableTiming), "hipEventCreate");
            check(hipEventCreateWithFlags(&done[1], hipEventDisableTiming), "hipEventCreate");
            Ctrl* host_slots = nullptr;
            check(hipHostMalloc(reinterpret_cast<void**>(&host_slots), 2 * sizeof(Ctrl), 0),
                  "hipHostMalloc");
            auto launch_one = [&](int slot) {
                check(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch");
                launched += kSteps;
                check(hipMemcpyAsync(&host_slots[slot], ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost,
                                     stream),
                      "hipMemcpyAsync(ctrl)");
                check(hipEventRecord(done[slot], stream), "hipEventRecord");
            };
            try {
                // steps whose count is fixed need no polling beyond the launch count
                const long long fixed = allow_max_steps
                                            ? (static_cast<long long>(max_steps) + kSteps - 1) / kSteps
                                            : -1;
                int slot = 0;
                long long issued = 0;
                int seen_steps = -1;  // step count after the previous replay read back
                launch_one(slot);
                ++issued;
                while (true) {
                    if (fixed >= 0 && issued >= fixed) {
                        break;
                    }
                    launch_one(slot ^ 1);
                    ++issued;
                    check(hipEventSynchronize(done[slot]), "hipEventSynchronize");
                    const Ctrl& c = host_slots[slot];
                    // parity after the replay that filled this slot is 0 (kSteps is even)
                    const bool finished =
                        c.bad_dt != 0 || !(c.t[0] < t_end && c.steps[0] < max_steps);
                    if (finished && fixed < 0) {
                        break;
                    }
                    // an unfinished solve whose replay took no step would replay forever
                    if (!finished && c.steps[0] <= seen_steps) {
                        throw std::runtime_error("solve: a replay of steps made no progress");
                    }
                    seen_steps = c.steps[0];
                    slot ^= 1;
                }
                check(hipStreamSynchronize(stream), "hipStreamSynchronize");
            } catch (...) {
                (void)hipStreamSynchronize(stream);
                (void)hipHostFree(host_slots);
                (void)hipEventDestroy(done[0]);
                (void)hipEventDestroy(done[1]);
                throw;
            }
            (void)hipHostFree(host_slots);
            (void)hipEventDestroy(done[0]);
            (void)hipEventDestroy(done[1]);
        }
        // final ghost + wall fill: include/mhd/time_integrator.hpp:101-104
        const Gate open{ctrl, 0.0, 0, -1};
        check(mhdhip::launch_ghost_fill<R>(u, geom, open, stream), "ghost_fill");
        if (masked && walls) {
            check(mhdhip::launch_wall_fill<R>(u, geom, plan_dev, plan_count, p.gamma, open, stream),
                  "wall_fill");
        }
        check(hipEventRecord(ev_stop, stream), "hipEventRecord");
        check(hipMemcpyAsync(ctrl_host, ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(ctrl)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ev_start, ev_stop), "hipEventElapsedTime");
        last_seconds = static_cast<double>(ms) * 1e-3;

        const int q = static_cast<int>(launched & 1);
        mhd_hip_result r{};
        r.t = ctrl_host->t[q];
        r.steps = ctrl_host->steps[q];
        r.c_h = ctrl_host->ch_last[q];
        dt_count = std::min(r.steps, dt_cap);
        if (ctrl_host->bad_dt != 0) {
            throw InvalidArgument("ssp_rk2_step requires positive dt");
        }
        if (!allow_max_steps && r.t < t_end && r.steps >= max_steps) {
            throw std::runtime_error("solve: reached max_steps before t_end");
        }
        return r;
    }

    mhd_hip_result solve(const mhd_hip_params& p)
    {
        switch (precision) {
            case MHD_HIP_FP64: return run<double, 0>(p, p.t_end, p.max_steps, false);
            case MHD_HIP_FP64_FAST: return run<double, 1>(p, p.t_end, p.max_steps, false);
            default: return run<float, 0>(p, p.t_end, p.max_steps, false);
        }
    }

    mhd_hip_result advance(const mhd_hip_params& p, int steps)
    {
        if (steps < 0) {
            throw InvalidArgument("steps must be non-negative");
        }
        const double inf = std::numeric_limits<double>::infinity();
        switch (precision) {
            case MHD_HIP_FP64: return run<double, 0>(p, inf, steps, true);
            case MHD_HIP_FP64_FAST: return run<double, 1>(p, inf, steps, true);
            default: return run<float, 0>(p, inf, steps, true);
        }
    }

    template <typename R>
    void fill_boundaries_t(double gamma, int what)
    {
        const Gate open{ctrl, 0.0, 0, -1};
        R* u = static_cast<R*>(U);
        if (what & MHD_HIP_FILL_GHOSTS) {
            check(mhdhip::launch_ghost_fill<R>(u, geom, open, stream), "ghost_fill");
        }
        if ((what & MHD_HIP_FILL_WALLS) && masked && walls) {
            check(mhdhip::launch_wall_fill<R>(u, geom, plan_dev, plan_count, gamma, open, stream),
                  "wall_fill");
        }
    }

    void fill_boundaries(double gamma, int what)
    {
        check(hipSetDevice(device), "hipSetDevice");
        if (precision != MHD_HIP_FP32) {
            fill_boundaries_t<double>(gamma, what);
        } else {
            fill_boundaries_t<float>(gamma, what);
        }
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // totals() with its partial sums and output in the process-wide `scratch` array of this
    // device instead of buffers the solver allocates: no allocation on any call. Calls from
    // this process are serialised (one at a time per process, all devices).
    void totals_noalloc(double out[9])
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mutex);
        char* base = static_cast<char*>(noalloc_scratch(device));
        check(hipSetDevice(device), "hipSetDevice");
        auto* partial = reinterpret_cast<mhdhip::DD*>(base);
        auto* dev_out = reinterpret_cast<double*>(base + kNoallocPartials * sizeof(mhdhip::DD));
        if (precision != MHD_HIP_FP32) {
            check(mhdhip::launch_totals<double>(static_cast<double*>(U), geom, partial, dev_out,
                                                stream),
                  "totals");
        } else {
            check(mhdhip::launch_totals<float>(static_cast<float*>(U), geom, partial, dev_out,
                                               stream),
                  "totals");
        }
        check(hipMemcpyAsync(out, dev_out, 9 * sizeof(double), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(totals)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    void totals(double out[9])
    {
        check(hipSetDevice(device), "hipSetDevice");
        if (!totals_partial) {
            check(hipMalloc(reinterpret_cast<void**>(&totals_partial),
                            sizeof(mhdhip::DD) * 9 * mhdhip::kTotalsBlocks),
                  "hipMalloc(totals)");
            check(hipMalloc(reinterpret_cast<void**>(&totals_out), 9 * sizeof(double)),
                  "hipMalloc(totals)");
        }
        if (precision != MHD_HIP_FP32) {
            check(mhdhip::launch_totals<double>(static_cast<double*>(U), geom, totals_partial,
                                                totals_out, stream),
                  "totals");
        } else {
            check(mhdhip::launch_totals<float>(static_cast<float*>(U), geom, totals_partial,
                                               totals_out, stream),
                  "totals");
        }
        check(hipMemcpyAsync(out, totals_out, 9 * sizeof(double), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(totals)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    // Relative L1 distance per component of this solver's device state from `ref`'s (the
    // reference), over the interior, on this solver's stream after `ref`'s pending work.
    void rel_l1(mhd_hip_solver& ref, double out[9])
    {
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
        check(mhdhip::launch_rel_l1_any(U, precision != MHD_HIP_FP32, ref.U,
                                        ref.precision != MHD_HIP_FP32, geom, l1_partial,
                                        totals_out, stream),
              "rel_l1");
        check(hipMemcpyAsync(out, totals_out, 9 * sizeof(double), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(rel_l1)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

    double max_div_b()
    {
        check(hipSetDevice(device), "hipSetDevice");
        if (!totals_out) {
            check(hipMalloc(reinterpret_cast<void**>(&totals_out), 9 * sizeof(double)),
                  "hipMalloc(totals)");
        }
        // include/grid/grid.hpp:24-25 (dx, dy), inititial_condition.hpp:276-277 (0.5 / d)
        const double dx = (grid.x_max - grid.x_min) / static_cast<double>(grid.nx);
        const double dy = (grid.y_max - grid.y_min) / static_cast<double>(grid.ny);
        auto* out = reinterpret_cast<unsigned long long*>(totals_out);
        if (precision != MHD_HIP_FP32) {
            check(mhdhip::launch_max_div_b<double>(static_cast<double*>(U), geom, 0.5 / dx,
                                                   0.5 / dy, out, stream),
                  "max_div_b");
        } else {
            check(mhdhip::launch_max_div_b<float>(static_cast<float*>(U), geom, 0.5 / dx,
                                                  0.5 / dy, out, stream),
                  "max_div_b");
        }
        double v = 0.0;
        check(hipMemcpyAsync(&v, totals_out, sizeof(double), hipMemcpyDeviceToHost, stream),
              "hipMemcpyAsync(max_div_b)");
        check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        return v;
    }

    void set_dt_log(int cap)
    {
        if (cap < 0) {
            throw InvalidArgument("dt log capacity must be non-negative");
        }
        check(hipSetDevice(device), "hipSetDevice");
        if (dt_log) {
            (void)hipFree(dt_log);
            dt_log = nullptr;
        }
        dt_cap = cap;
        dt_count = 0;
        if (cap > 0) {
            check(hipMalloc(reinterpret_cast<void**>(&dt_log), static_cast<size_t>(cap) * sizeof(double)),
                  "hipMalloc(dt_log)");
        }
        invalidate_graph();
    }
};

namespace mhdhip {

int default_seg_rows(int nx, int ny)
{
    // aim for ~8192 waves; segments of at least 8 rows amortise the 4-row halo
    const long long strips = (nx + kStripCells - 1) / kStripCells;
    long long rows = strips * ny / 8192;
    int r = 8;
    while (r < rows && r < 128) {
        r *= 2;
    }
    return std::min(r, std::max(1, ny));
}

Tiling make_tiling(int nx, int ny, int seg_rows)
{
    Tiling t{};
    t.strips = (nx + kStripCells - 1) / kStripCells;
    t.seg_rows = std::max(1, seg_rows);
    t.segs = (ny + t.seg_rows - 1) / t.seg_rows;
    return t;
}

}  // namespace mhdhip

extern "C" {

int mhd_hip_available(void)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
        return 0;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        return 0;
    }
    return device_is_gfx950(dev) ? 1 : 0;
}

const char* mhd_hip_last_error(void)
{
    return mhdhip::last_error_slot().c_str();
}

int mhd_hip_create(const mhd_hip_grid* grid, int precision, mhd_hip_solver** out)
{
    return guarded([&] {
        if (grid == nullptr || out == nullptr) {
            throw InvalidArgument("null argument");
        }
        *out = nullptr;
        auto s = std::make_unique<mhd_hip_solver>();
        s->device = -1;
        s->init(*grid, precision);
        *out = s.release();
    });
}

void mhd_hip_destroy(mhd_hip_solver* s)
{
    delete s;
}

int mhd_hip_device_name(const mhd_hip_solver* s, char* buf, size_t len)
{
    return guarded([&] {
        if (buf == nullptr || len == 0) {
            throw InvalidArgument("null buffer");
        }
        int dev = 0;
        if (s != nullptr) {
            dev = s->device;
        } else {
            check(hipGetDevice(&dev), "hipGetDevice");
        }
        hipDeviceProp_t prop;
        check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        // the marketing name can be empty (some drivers and containers): fall back to the
        // device index so the string still names one device
        if (prop.name[0] != '\0') {
            std::snprintf(buf, len, "HIP: %s (%s)", prop.name, prop.gcnArchName);
        } else {
            std::snprintf(buf, len, "HIP: device %d (%s)", dev, prop.gcnArchName);
        }
    });
}

int mhd_hip_precision(const mhd_hip_solver* s)
{
    return s ? s->precision : -1;
}

int mhd_hip_configure(mhd_hip_solver* s, int bc_x, int bc_y, const uint8_t* mask, size_t mask_len,
                      int reflective_walls)
{
    return guarded([&] {
        if (!s) throw InvalidArgument("null solver");
        s->configure(bc_x, bc_y, mask, mask_len, reflective_walls);
    });
}

int mhd_hip_upload(mhd_hip_solver* s, const double* const comps[9], size_t n)
{
    return guarded([&] {
        if (!s || !comps) throw InvalidArgument("null argument");
        for (int c = 0; c < 9; ++c) {
            if (!comps[c]) throw InvalidArgument("null component pointer");
        }
        s->upload(comps, n);
    });
}

int mhd_hip_download(mhd_hip_solver* s, double* const comps[9], size_t n)
{
    return guarded([&] {
        if (!s || !comps) throw InvalidArgument("null argument");
        for (int c = 0; c < 9; ++c) {
            if (!comps[c]) throw InvalidArgument("null component pointer");
        }
        s->download(comps, n);
    });
}

int mhd_hip_solve(mhd_hip_solver* s, const mhd_hip_params* params, mhd_hip_result* result)
{
    return guarded([&] {
        if (!s || !params || !result) throw InvalidArgument("null argument");
        *result = s->solve(*params);
    });
}

int mhd_hip_advance_steps(mhd_hip_solver* s, const mhd_hip_params* params, int steps,
                          mhd_hip_result* result)
{
    return guarded([&] {
        if (!s || !params || !result) throw InvalidArgument("null argument");
        *result = s->advance(*params, steps);
    });
}

int mhd_hip_set_dt_log(mhd_hip_solver* s, int capacity)
{
    return guarded([&] {
        if (!s) throw InvalidArgument("null solver");
        s->set_dt_log(capacity);
    });
}

int mhd_hip_dt_log(mhd_hip_solver* s, double* out, int capacity, int* count)
{
    return guarded([&] {
        if (!s || !count) throw InvalidArgument("null argument");
        const int n = std::min(capacity, s->dt_count);
        *count = std::max(0, n);
        if (n > 0) {
            if (!out) throw InvalidArgument("null output buffer");
            check(hipSetDevice(s->device), "hipSetDevice");
            s->copy_sync(out, s->dt_log, static_cast<size_t>(n) * sizeof(double),
                         hipMemcpyDeviceToHost, "hipMemcpy(dt_log)");
        }
    });
}

int mhd_hip_fill_boundaries(mhd_hip_solver* s, double gamma, int what)
{
    return guarded([&] {
        if (!s) throw InvalidArgument("null solver");
        s->fill_boundaries(gamma, what);
    });
}

int mhd_hip_totals(mhd_hip_solver* s, double sums[9])
{
    return guarded([&] {
        if (!s || !sums) throw InvalidArgument("null argument");
        s->totals(sums);
    });
}

int mhd_hip_totals_noalloc(mhd_hip_solver* s, double sums[9])
{
    return guarded([&] {
        if (!s || !sums) throw InvalidArgument("null argument");
        s->totals_noalloc(sums);
    });
}

int mhd_hip_rel_l1(mhd_hip_solver* a, mhd_hip_solver* ref, double l1[9])
{
    return guarded([&] {
        if (!a || !ref || !l1) throw InvalidArgument("null argument");
        if (a->grid.nx != ref->grid.nx || a->grid.ny != ref->grid.ny ||
            a->grid.ng != ref->grid.ng) {
            throw InvalidArgument("rel_l1: the two states must have the same grid");
        }
        if (a->device != ref->device) {
            throw InvalidArgument("rel_l1: the two solvers are on different devices");
        }
        a->rel_l1(*ref, l1);
    });
}

int mhd_hip_max_div_b(mhd_hip_solver* s, double* value)
{
    return guarded([&] {
        if (!s || !value) throw InvalidArgument("null argument");
        *value = s->max_div_b();
    });
}

int mhd_hip_set_seg_rows(mhd_hip_solver* s, int rows)
{
    return guarded([&] {
        if (!s) throw InvalidArgument("null solver");
        if (rows < 0) throw InvalidArgument("rows per wave segment must be >= 0");
        s->seg_rows = rows;
        s->invalidate_graph();
    });
}

int mhd_hip_get_seg_rows(mhd_hip_solver* s, int* rows)
{
    return guarded([&] {
        if (!s || !rows) throw InvalidArgument("null argument");
        *rows = s->last_rows;
    });
}

int mhd_hip_set_kernel_timing(mhd_hip_solver* s, int enable)
{
    return guarded([&] {
        if (!s) throw InvalidArgument("null solver");
        s->kernel_timing = enable != 0;
        s->stage_ms[0] = s->stage_ms[1] = 0.0;
        s->stage_launches[0] = s->stage_launches[1] = 0;
    });
}

int mhd_hip_kernel_times(const mhd_hip_solver* s, double* total_ms, long long* launches)
{
    return guarded([&] {
        if (!s || !total_ms || !launches) throw InvalidArgument("null argument");
        for (int k = 0; k < 2; ++k) {
            total_ms[k] = s->stage_ms[k];
            launches[k] = s->stage_launches[k];
        }
    });
}

int mhd_hip_last_solve_seconds(const mhd_hip_solver* s, double* seconds)
{
    return guarded([&] {
        if (!s || !seconds) throw InvalidArgument("null argument");
        *seconds = s->last_seconds;
    });
}

int mhd_hip_init_ic(const char* name, const mhd_hip_grid* grid, double gamma, double b0,
                    uint64_t seed, double* const comps[9])
{
    return guarded([&] {
        if (!grid || !comps) throw InvalidArgument("null argument");
        if (grid->nx < 1 || grid->ny < 1 || grid->ng < 0) throw InvalidArgument("bad grid");
        for (int c = 0; c < 9; ++c) {
            if (!comps[c]) throw InvalidArgument("null component pointer");
        }
        mhdhip::init_ic(name, *grid, gamma, b0, seed, comps);
    });
}

int mhd_hip_wall_plan(const mhd_hip_grid* grid, const uint8_t* mask, size_t mask_len,
                      int64_t* plan, size_t plan_capacity, size_t* count)
{
    return guarded([&] {
        if (!grid || !mask || !count) throw InvalidArgument("null argument");
        if (mask_len != static_cast<size_t>(grid->nx) * grid->ny) {
            throw InvalidArgument("FluidMask size does not match Grid2D interior");
        }
        std::vector<long long> p;
        const size_t n = mhdhip::wall_plan(*grid, mask, p);
        *count = n;
        if (plan != nullptr) {
            if (plan_capacity < p.size()) throw InvalidArgument("plan buffer too small");
            for (size_t k = 0; k < p.size(); ++k) {
                plan[k] = p[k];
            }
        }
    });
}

}  // extern "C"
