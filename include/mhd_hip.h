/*
 * mhd_hip.h - C ABI of the MI355X (gfx950) "hip" backend of the 2-D GLM-MHD solver.
 *
 * Drop-in boundary: every entry point below replaces one member of the reference's device
 * backend interface (include/mhd/gpu/device_solver.hpp:57-75) or its factory/discovery helpers
 * (device_solver.hpp:26-111), with plain pointers, sizes and integer status codes so that a C++
 * DeviceSolver adapter, a pybind11 module or ctypes can bind it. Paths cited are relative to the
 * reference repository root.
 *
 * Status codes: MHD_HIP_OK on success; MHD_HIP_EINVAL where the reference throws
 * std::invalid_argument; MHD_HIP_ERUNTIME where it throws std::runtime_error (device failure,
 * max_steps reached before t_end); MHD_HIP_ENODEV when no usable GPU is present.
 * mhd_hip_last_error() returns the calling thread's message for the last failure.
 *
 * State layout on the host side of upload/download: nine arrays (rho, mx, my, mz, energy, bx,
 * by, bz, psi), each (ny + 2 ng) * (nx + 2 ng) doubles, row-major, ghosts included
 * (include/state/state_field.hpp:9-45).
 */
#ifndef MHD_HIP_H
#define MHD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHD_HIP_ABI_VERSION 1

enum mhd_hip_status {
    MHD_HIP_OK = 0,
    MHD_HIP_EINVAL = 1,
    MHD_HIP_ERUNTIME = 2,
    MHD_HIP_ENODEV = 3,
    MHD_HIP_EFULL = 4 /* batch capacity (HBM budget or instance slots) exhausted; retry later */
};

/* mhd::Precision (device_solver.hpp:28) */
/* MHD_HIP_FP64: fp64, bit-identical to the reference's CPU path. MHD_HIP_FP32: FP32 storage
 * and arithmetic with fp64 pressure recovery and fast-speed discriminant. MHD_HIP_FP64_FAST
 * (no counterpart in the reference's Precision): fp64 storage and arithmetic with contracted
 * multiply-adds and Newton-refined hardware reciprocal and square root in the stepping
 * kernels; within 1e-12 relative L1 of MHD_HIP_FP64 at the same dt sequence, not bit-equal.
 * mhd_hip_precision reports it as created. */
enum mhd_hip_precision { MHD_HIP_FP64 = 0, MHD_HIP_FP32 = 1, MHD_HIP_FP64_FAST = 2 };

/* BoundaryConditionType (include/boundary_conditions/boundary_conditions.hpp:8) */
enum mhd_hip_bc { MHD_HIP_PERIODIC = 0, MHD_HIP_REFLECTIVE = 1, MHD_HIP_OUTFLOW = 2 };

/* SlopeLimiter (include/mhd/mhd_types.hpp:44) */
enum mhd_hip_limiter { MHD_HIP_MINMOD = 0, MHD_HIP_MC = 1 };

/* Grid2D (include/grid/grid.hpp:5-30): interior nx x ny, ng ghosts per side. */
typedef struct mhd_hip_grid {
    int nx, ny, ng;
    double x_min, x_max, y_min, y_max;
} mhd_hip_grid;

/* SolveParams (include/mhd/solve_types.hpp:8-16) with NonIdealConfig flattened
 * (include/physics_config/mhd_config.hpp:13-24). */
typedef struct mhd_hip_params {
    double t_end;
    double cfl;
    double gamma;
    double glm_alpha;
    int limiter;
    int ohmic, hall, ambipolar;
    double eta_ohm, eta_hall, eta_ambipolar;
    int max_steps;
} mhd_hip_params;

/* SolveResult (include/mhd/solve_types.hpp:18-22) */
typedef struct mhd_hip_result {
    double t;
    int steps;
    double c_h;
} mhd_hip_result;

typedef struct mhd_hip_solver mhd_hip_solver;

/* Replaces cuda_available()/metal_available() (device_solver.hpp:77-78): 1 if a gfx950 device
 * is present and the kernels load, else 0. Never fails. */
int mhd_hip_available(void);

/* Thread-local message of the last failing call on this thread ("" if none). */
const char* mhd_hip_last_error(void);

/* Replaces make_cuda_solver(grid, precision) (device_solver.hpp:81-82, cuda_solver.cu:786-798):
 * allocates the device state for `grid` on the current HIP device. EINVAL for ng < 2 or
 * nx, ny < 1; ENODEV without a device. */
int mhd_hip_create(const mhd_hip_grid* grid, int precision, mhd_hip_solver** out);

/* Replaces ~DeviceSolver. Null is accepted. */
void mhd_hip_destroy(mhd_hip_solver* s);

/* Replaces DeviceSolver::device_name() (device_solver.hpp:62). Writes a NUL-terminated name. */
int mhd_hip_device_name(const mhd_hip_solver* s, char* buf, size_t len);

/* Replaces DeviceSolver::precision() (device_solver.hpp:63). */
int mhd_hip_precision(const mhd_hip_solver* s);

/* Replaces DeviceSolver::configure(bc, mask, reflective_walls) (device_solver.hpp:66-67).
 * mask: ny * nx bytes, row-major over the interior, 1 = fluid (include/mhd/mask.hpp:17-20);
 * NULL (or all fluid) means no solids. EINVAL for Reflective boundaries. */
int mhd_hip_configure(mhd_hip_solver* s, int bc_x, int bc_y, const uint8_t* mask, size_t mask_len,
                      int reflective_walls);

/* Replaces DeviceSolver::upload(const StateField&) (device_solver.hpp:70). comps[c] points at
 * plane c of the full state; n must equal (nx + 2 ng) * (ny + 2 ng). */
int mhd_hip_upload(mhd_hip_solver* s, const double* const comps[9], size_t n);

/* Replaces DeviceSolver::download(StateField&) (device_solver.hpp:71). */
int mhd_hip_download(mhd_hip_solver* s, double* const comps[9], size_t n);

/* Replaces DeviceSolver::solve(const SolveParams&) (device_solver.hpp:74): advances the device
 * state by the relative time params->t_end with the contract of ::solve()
 * (include/mhd/time_integrator.hpp:62-106): EINVAL for t_end < 0 or cfl <= 0, ERUNTIME when
 * max_steps is reached before t_end; ghosts and solid cells are filled at the end. */
int mhd_hip_solve(mhd_hip_solver* s, const mhd_hip_params* params, mhd_hip_result* result);

/* Fixed step count (benchmarks, dt-sequence checks): exactly `steps` steps of the scheme, as
 * ::solve() with t_end = +inf and max_steps = steps, without the max_steps error. The
 * reference's solve loop (time_integrator.hpp:84-95) with the loop bound as the only change. */
int mhd_hip_advance_steps(mhd_hip_solver* s, const mhd_hip_params* params, int steps,
                          mhd_hip_result* result);

/* Records the dt of the first `capacity` steps of each later solve/advance call
 * (capacity 0 disables). mhd_hip_dt_log copies min(count, capacity) values and returns the
 * number copied through *count. */
int mhd_hip_set_dt_log(mhd_hip_solver* s, int capacity);
int mhd_hip_dt_log(mhd_hip_solver* s, double* out, int capacity, int* count);

/* Ghost fill and/or solid-wall fill of the device state, as MHDSolver::initialize and set_mask
 * do on the host state (include/mhd/solver_api.hpp:148-154, :286-289). `what` is a bit set:
 * MHD_HIP_FILL_GHOSTS, MHD_HIP_FILL_WALLS (walls only with an active mask and reflective walls);
 * ghosts are filled before walls. */
#define MHD_HIP_FILL_GHOSTS 1
#define MHD_HIP_FILL_WALLS 2
int mhd_hip_fill_boundaries(mhd_hip_solver* s, double gamma, int what);

/* Totals of the nine conserved components over the interior cells of the device state
 * (multiply by dx*dy for integrals), summed on the device in double-double arithmetic in a
 * fixed order: ~106-bit sums whatever the storage precision, identical on every call. Used for
 * mass/energy drift; the reference sums plainly on the host (diagnostics.hpp:42-57). */
int mhd_hip_totals(mhd_hip_solver* s, double sums[9]);

/* mhd_hip_totals without any allocation, on any call. No counterpart in the reference: for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). The same kernels, hence the same bits, with the partial sums and the totals in
 * a static device array of the library (73,800 bytes per device, held until the process
 * ends) instead of buffers the solver allocates on its first mhd_hip_totals call. The array
 * holds the partials of a fixed number of workgroups whatever the grid, so no grid is too
 * large for it; devices beyond index 63 are rejected (EINVAL). Every call on a device shares
 * that device's array (each process holds its own copy), so calls that overlapped on one
 * device would race: the library serialises them within the process. */
int mhd_hip_totals_noalloc(mhd_hip_solver* s, double sums[9]);

/* max |dBx/dx + dBy/dy| over the interior by central differences of the device state, ghosts
 * included, in fp64 with the reference's operation order and max rule; after a solve (which
 * ends with the ghost and wall fill) it equals the reference's MHDSolver::max_div_b
 * (solver_api.hpp:344-347, inititial_condition.hpp:273-287) for the same state. */
int mhd_hip_max_div_b(mhd_hip_solver* s, double* value);

/* Relative L1 distance of one device state from another of the same grid, per conserved
 * component over the interior, computed on the device: l1[c] = sum |a - ref| / sum |ref| (the
 * absolute sum where sum |ref| is 0). The states may differ in precision (FP32 against fp64);
 * an oracle state is compared by uploading it into a second solver. Double-double sums in a
 * fixed order: the same bits on every call. Both solvers must be on one device.
 * No counterpart call in the reference, which compares downloaded states on the host
 * (tools/mhd_backend_compare.cpp:169-207); the per-field relative L1 is its accuracy metric. */
int mhd_hip_rel_l1(mhd_hip_solver* a, mhd_hip_solver* ref, double l1[9]);

/* Seconds of device time spent in the last solve/advance call (HIP events on the solver's
 * stream around the step loop, excluding upload/download). */
int mhd_hip_last_solve_seconds(const mhd_hip_solver* s, double* seconds);

/* Rows of the grid marched by one wavefront of the stage kernels (the tiling's only free
 * parameter). 0 (default) selects the autotuned value: on the first solve of a shape the pair
 * of stage kernels of one step is timed for candidate lengths and the fastest is kept for the
 * process; a positive integer in the environment variable MHD_HIP_SEG_ROWS replaces the
 * tuning for every shape. Results do not depend on this value. mhd_hip_get_seg_rows reports
 * the value last used. */
int mhd_hip_set_seg_rows(mhd_hip_solver* s, int rows);
int mhd_hip_get_seg_rows(mhd_hip_solver* s, int* rows);

/* Kernel timing for measurement: while enabled, fixed-step calls (mhd_hip_advance_steps) launch
 * their steps eagerly (no graph) with a HIP event pair on the solver's stream around each
 * stage kernel. mhd_hip_kernel_times returns, for the stage-1 and stage-2 kernels, the summed
 * event time in ms and the number of launches timed since timing was (re)enabled. */
int mhd_hip_set_kernel_timing(mhd_hip_solver* s, int enable);
int mhd_hip_kernel_times(const mhd_hip_solver* s, double total_ms[2], long long launches[2]);

/* Host-side initial condition (include/initial_condition/inititial_condition.hpp:28-270):
 * writes the conserved interior of the nine planes for a reference IC name (uniform, sine,
 * blast, orszag_tang, sod, brio_wu, rotor, kelvin_helmholtz / kh, alfven / alfven_wave, with
 * the reference aliases) or one of this backend's parametrised workloads:
 *   "orszag_tang_b0"   Orszag-Tang with B scaled by `b0` (b0 = 1/sqrt(4 pi) for the standard
 *                      normalisation)
 *   "rotor_b0"         rotor with Bx = b0 * 5/sqrt(4 pi)
 *   "kh_tanh"          tanh shear layer, vx = +/-0.5, width 0.05, Bx = 0.1, vy seeded noise
 *                      of amplitude 1e-2 from `seed`
 *   "alfven_oblique"   circularly polarised Alfven wave, amplitude 0.1, along angle atan(2)
 * `b0` and `seed` are ignored by the reference names. Ghost cells are not written.
 * EINVAL for an unknown name. Runs on the CPU; needs no device. */
int mhd_hip_init_ic(const char* name, const mhd_hip_grid* grid, double gamma, double b0,
                    uint64_t seed, double* const comps[9]);

/* Host-side solid-wall plan (include/mhd/mask.hpp:85-149): for each solid interior cell in
 * row-major order, (destination cell, first fluid interior neighbour in the order +x, -x, +y,
 * -y or -1, axis 0/1). plan must hold 3 * (number of solid cells) entries; returns the count
 * through *count. Runs on the CPU; needs no device. */
int mhd_hip_wall_plan(const mhd_hip_grid* grid, const uint8_t* mask, size_t mask_len,
                      int64_t* plan, size_t plan_capacity, size_t* count);

/* ---- Batched ensemble engine ------------------------------------------------------------
 * No counterpart in the reference, which advances one grid per DeviceSolver
 * (device_solver.hpp:57-75) and runs concurrent solves only on separate host threads
 * (docs/performance.md:80-81). Many independent grids of different sizes stay resident in an
 * HBM arena of a caller-given budget and advance together in one persistent kernel per RK
 * stage, each with its own dt (device-side CFL per instance), gamma, CFL number, end time and
 * periodic/outflow boundaries; limiter and non-ideal terms are shared by the batch. Each
 * instance follows the ::solve() contract (time_integrator.hpp:62-106) exactly as a single-run
 * solver would: its result does not depend on the other instances. */

typedef struct mhd_hip_batch mhd_hip_batch;

/* One solve: grid, boundaries and SolveParams fields that may differ per instance. */
typedef struct mhd_hip_instance {
    int nx, ny; /* interior cells; ng is 2 */
    double x_min, x_max, y_min, y_max;
    int bc_x, bc_y; /* MHD_HIP_PERIODIC or MHD_HIP_OUTFLOW */
    double t_end;   /* relative end time; +inf with max_steps for a fixed step count */
    double cfl, gamma, glm_alpha;
    int max_steps;
} mhd_hip_instance;

typedef struct mhd_hip_batch_stats {
    int instances; /* admitted (running or finished, not yet removed) */
    int active;    /* still stepping (unfinished, not paused) */
    int tiles;     /* wave work items in the current tile list */
    int seg_rows;  /* rows per tile of the current tile list (0: capacity fill) */
    int blocks;    /* workgroups of the largest persistent stage kernel */
    size_t bytes_used, bytes_budget;
    double last_seconds; /* device time of the last mhd_hip_batch_advance */
    int paused;          /* resident, unfinished, excluded from stepping */
} mhd_hip_batch_stats;

/* HBM bytes one instance of an nx x ny grid occupies in the arena (U and U*; an instance with
 * solid cells also holds its mask, nx * ny bytes, and 24 bytes per solid cell of wall plan,
 * each rounded up to 256 bytes). */
size_t mhd_hip_batch_instance_bytes(int precision, int nx, int ny);

/* physics: the limiter and non-ideal fields are the scheme of the instances admitted with
 * mhd_hip_batch_add; others ignored. */
int mhd_hip_batch_create(int precision, const mhd_hip_params* physics, size_t hbm_budget_bytes,
                         int max_instances, mhd_hip_batch** out);
void mhd_hip_batch_destroy(mhd_hip_batch* b);

/* Admission: copies the full fp64 state (nine planes of (ny+4)*(nx+4), ghosts included) into
 * the arena and runs the instance's first CFL reduction. MHD_HIP_EFULL when the HBM budget or
 * the instance slots are exhausted. Instances join at the next replay. */
int mhd_hip_batch_add(mhd_hip_batch* b, const mhd_hip_instance* inst, const double* const comps[9],
                      int* id);

/* Admission with the instance's own scheme: its slope limiter and non-ideal terms are taken
 * from `scheme` (limiter, ohmic/hall/ambipolar and their coefficients; the other fields of
 * `scheme` are ignored, `inst` gives them), so one engine advances instances of different
 * schemes together. `mask` (NULL: all fluid) is the ny x nx interior FluidMask of
 * include/mhd/mask.hpp:17-83 (1 = fluid, `mask_len` = nx * ny, else EINVAL): the CFL speed is
 * taken over fluid cells only and the right-hand side is zero on solid cells; with
 * `reflective_walls` the solid cells take the reflective wall fill (include/mhd/mask.hpp:
 * 85-149) after the ghost fill before every stage and at the end, as MHDSolver does with
 * set_walls_reflective(true). The mask and the wall plan are kept in the arena beside the
 * state. Otherwise as mhd_hip_batch_add. */
int mhd_hip_batch_add_config(mhd_hip_batch* b, const mhd_hip_instance* inst,
                             const mhd_hip_params* scheme, const uint8_t* mask, size_t mask_len,
                             int reflective_walls, const double* const comps[9], int* id);

/* Releases an instance's slot and memory (finished or not). */
int mhd_hip_batch_remove(mhd_hip_batch* b, int id);

/* Replays the step graph until no instance is active, or at most max_replays times
 * (max_replays < 0: no limit). *active receives the number of instances still stepping.
 * MHD_HIP_ERUNTIME if a replay advanced none of the instances it stepped (a replay that makes
 * no progress would repeat forever). */
int mhd_hip_batch_advance(mhd_hip_batch* b, int max_replays, int* active);

/* Does before the next replay what its launch needs (no counterpart in the reference, whose
 * solver plans nothing across calls): reads the control blocks, schedules the stepping set,
 * makes the timed row choice if it is pending, builds the tile tables and captures the step
 * graph. mhd_hip_batch_advance does this itself when needed; a caller that admitted a burst of
 * instances can call this first so that the planning is not part of the next advance. */
int mhd_hip_batch_prepare(mhd_hip_batch* b);

/* Result so far (t, steps, c_h as SolveResult) and whether the instance has finished; the
 * ::solve() errors are reported here (EINVAL for dt <= 0, ERUNTIME for max_steps before t_end). */
int mhd_hip_batch_result(mhd_hip_batch* b, int id, mhd_hip_result* result, int* finished);

/* Ghost fill (the solve's final fill) and copy of the instance's state to nine fp64 planes. */
int mhd_hip_batch_download(mhd_hip_batch* b, int id, double* const comps[9], size_t n);

/* Scheduling hooks. A paused instance stays resident (state and control block untouched) but
 * is left out of the step graph's tile list from the next replay on; mhd_hip_batch_advance
 * returns once no unpaused instance is stepping. mhd_hip_batch_progress reports t and steps
 * as of the last replay and an estimate of the instance's dt (t / steps, or before its first
 * step the dt of its admission CFL reduction). */
int mhd_hip_batch_set_paused(mhd_hip_batch* b, int id, int paused);
int mhd_hip_batch_progress(mhd_hip_batch* b, int id, double* t, int* steps, double* dt_estimate);

/* Engine-side scheduling, applied before every replay. MHD_HIP_SCHED_ALL steps every
 * unfinished, unpaused instance (the default). MHD_HIP_SCHED_SRPT orders the unfinished
 * instances by estimated remaining work (interior cells x remaining steps, the steps from the
 * dt estimate above and capped by max_steps) and steps the least-work ones until
 * work_cap_cells interior cells are stepping; it sets the paused flags itself, overriding
 * mhd_hip_batch_set_paused. Scheduling changes when a solve is computed, never its result. */
#define MHD_HIP_SCHED_ALL 0
#define MHD_HIP_SCHED_SRPT 1
int mhd_hip_batch_set_schedule(mhd_hip_batch* b, int policy, long long work_cap_cells);

/* Ids of the instances that have finished and are still resident (at most cap of them, in
 * slot order), as of the last replay; *count is the number written. */
int mhd_hip_batch_poll_finished(mhd_hip_batch* b, int* ids, int cap, int* count);

/* mhd_hip_max_div_b for one instance, after its final ghost fill (the state download sees). */
int mhd_hip_batch_max_div_b(mhd_hip_batch* b, int id, double* value);

/* mhd_hip_totals for one instance (interior, double-double, deterministic). */
int mhd_hip_batch_totals(mhd_hip_batch* b, int id, double sums[9]);

/* mhd_hip_rel_l1 for instance id_a of engine a against instance id_ref of engine ref (the
 * same engine or another on the same device, e.g. FP32 against fp64). */
int mhd_hip_batch_rel_l1(mhd_hip_batch* a, int id_a, mhd_hip_batch* ref, int id_ref,
                         double l1[9]);

/* Work partition of the batch kernel: tiles (instance, strip, row range) of fixed row
 * segments, in one queue per XCD from which the waves of the persistent grid take tiles as
 * they finish the last (largest instances first, band by band). 0 (default): the segment
 * length a load model picks whenever the stepping set changes (the costliest queue spread
 * over its XCD's waves plus half a tile; suits a set that changes, as in serving); -1: before
 * the first replay, up to three lengths the load model ranks within 12 % of its best are
 * timed on dry steps (stage 2 writes the scratch state, so no instance advances) and the
 * fastest is kept (suits a set that stays the same); > 0: segments of that many rows;
 * < -1: the capacity fill, every strip cut into row pieces of at least -rows rows poured
 * into per-wave lists in order up to a common capacity. Results do not depend on this value. */
int mhd_hip_batch_set_seg_rows(mhd_hip_batch* b, int rows);

int mhd_hip_batch_info(const mhd_hip_batch* b, mhd_hip_batch_stats* stats);

/* Measurement, as mhd_hip_set_kernel_timing: while enabled, each replay's steps are launched
 * eagerly with a HIP event pair around every stage kernel; totals per stage (ms, launches). */
int mhd_hip_batch_set_kernel_timing(mhd_hip_batch* b, int enable);
int mhd_hip_batch_kernel_times(const mhd_hip_batch* b, double total_ms[2], long long launches[2]);

/* Measured bandwidth of one RK stage's memory traffic without its arithmetic, on the current
 * device: a streaming kernel that reads 9 planes and writes 9 (stage 1) or reads 18 and writes
 * 9 in place (stage 2), `cells` elements per plane in fp64 or FP32 (precision), timed with HIP
 * events over `reps` launches after two untimed ones; *gbs = bytes moved / mean launch time.
 * Allocates and frees its own buffers (2 x 9 x cells elements). The bandwidth the memory
 * system delivers for the stage's mix of reads and writes, reported beside the published HBM
 * peak. No counterpart in the reference (a measurement aid). */
int mhd_hip_stream_bandwidth(int precision, long long cells, int stage, int reps, double* gbs);

#ifdef __cplusplus
}
#endif

#endif /* MHD_HIP_H */
