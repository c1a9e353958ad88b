"""Benchmark of the hip backend's hot path: SSP-RK2 steps of the GLM-MHD MUSCL/HLL stencil.

Headline workload (one GPU): the seeded mixed-size ensemble -- 512 independent Orszag-Tang /
rotor instances, square grids drawn uniformly from {64, 96, 128, 192, 256}, gamma ~ U[1.4, 5/3],
field scale ~ U[0.5, 2], minmod, HLL, CFL 0.4 -- advanced by the batched FP32-bulk engine (one
persistent kernel per RK stage, per-instance dt). A "step" is one SSP-RK2 step of every instance.
The states are resident in HBM before the timed region; admission/upload are outside it.

value: batched FP32 Mcell-updates/s = sum(N^2) * steps / seconds / 1e6 (the reference's
definition, tools/mhd_backend_compare.cpp:295,527), timed on the production path (graph replays,
bracketed by stream synchronisations); the same steps with HIP events around every stage kernel
give the kernel durations behind the roofline. Beside it: the same ensemble on the plain fp64 single-run kernel one instance at a time
(the reference's execution model) and the ratio; the per-field relative L1 of the FP32 results
against those fp64 results; the batched fp64 engine; the single large run (Orszag-Tang 4096^2)
in fp64 and FP32. roofline: the dominant stage kernel's algorithmic bytes over its average
duration. cpu_baseline: the reference's own CPU solver (Release flags) on a bounded sample of
the same ensemble on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--instances 512] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mcell-updates/s batched FP32 vs plain fp64 at L1<=1e-4; p50/p99 solve latency"
HBM_PEAK_GBS = 8000.0
B0_STD = 1.0 / math.sqrt(4.0 * math.pi)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---- single large run (Orszag-Tang N^2) ----------------------------------------------------

def make_device(m, nx, ny, precision, bc=(0, 0), y_max=1.0):
    from mhd_gpu_solver_amd import _lib
    grid = _lib.Grid(nx, ny, 2, 0.0, 1.0, 0.0, y_max)
    dev = m.DeviceSolver(grid, _lib.PRECISIONS[precision])
    dev.configure(bc[0], bc[1], None, True)
    return dev


def params(m, gamma=5.0 / 3.0):
    p = m.SolveParams()
    p.limiter = m.SlopeLimiter.Minmod
    p.cfl = 0.4
    p.gamma = gamma
    p.glm_alpha = 0.1
    return p


def single_leg(m, n, precision, steps, warmup, U0, ny=None, bc=(0, 0), y_max=1.0):
    """`steps` timed steps of the single-run solver after `warmup` (an n x n grid, or n x ny
    over [0, 1] x [0, y_max])."""
    ny = n if ny is None else ny
    dev = make_device(m, n, ny, precision, bc, y_max)
    t0 = time.perf_counter()
    dev.upload(U0)                                   # full fp64 host state, ghosts included
    upload_s = time.perf_counter() - t0
    dev.fill_boundaries(5.0 / 3.0)
    p = params(m)
    if warmup:
        dev.advance_steps(p, warmup)
    t0 = time.perf_counter()
    dev.advance_steps(p, steps)                      # production path: graph replays
    graph_s = time.perf_counter() - t0
    tot0 = dev.totals()
    dev.set_kernel_timing(True)                      # events around every stage kernel
    t0 = time.perf_counter()
    dev.advance_steps(p, steps)
    wall = time.perf_counter() - t0
    tot1 = dev.totals()
    drift = {"mass": float(f"{abs(tot1[0] - tot0[0]) / abs(tot0[0]):.3e}"),
             "energy": float(f"{abs(tot1[4] - tot0[4]) / abs(tot0[4]):.3e}")}
    k = dev.kernel_times()
    rows = dev.seg_rows()
    dev.set_kernel_timing(False)
    U = np.zeros_like(U0)
    t0 = time.perf_counter()
    dev.download(U)
    download_s = time.perf_counter() - t0
    if not np.isfinite(U).all():
        raise RuntimeError(f"{precision}: non-finite state after the timed steps")
    cells = n * ny
    # what a caller handing over host buffers sees for one call of `steps` steps: upload,
    # graph-path steps, download (MHDSolver.run moves the state once each way per call)
    pcie = {"mcups": round(cells * steps / (upload_s + graph_s + download_s) / 1e6, 2),
            "steps": steps, "upload_s": round(upload_s, 4), "download_s": round(download_s, 4)}
    return dict(mcups=cells * steps / graph_s / 1e6, ms_per_step=graph_s * 1e3 / steps,
                timed_mcups=cells * steps / wall / 1e6, kernels=k, seg_rows=rows,
                cells=cells, drift=drift, pcie=pcie)


def measured_traffic(kernel_prefix, elem_bytes, workload, kernel="k_stage", fast=0):
    """HBM bytes per launch of a stage kernel from the committed FETCH_SIZE/WRITE_SIZE summary
    (profiles/traffic_<kernel>_<workload>.json), or None when none exists for this workload.
    The kernel is the unmasked ideal one of the precision (and, for fp64, of the exact or the
    fast build: the last template argument)."""
    path = ROOT / "profiles" / f"traffic_{kernel}_{workload}.json"
    if not path.exists():
        return None
    data = json.loads(path.read_text())
    want = ("double" if elem_bytes == 8 else "float") + f", {fast}>"
    for name, v in data.get("kernels", {}).items():
        if name.startswith(kernel_prefix) and name.endswith(want):
            return v["traffic_bytes"]
    return None


def mix_ceiling(m, precision, cells):
    """Measured GB/s of a plain streaming kernel with each stage's read/write mix at this many
    cells per plane (no arithmetic): what the memory system delivers for that mix on this
    device, reported beside the published peak."""
    return {st: round(m.stream_bandwidth(precision, cells, st), 1) for st in (1, 2)}


def roofline(kernels, cells, ms_per_step, elem_bytes, name, traffic=None, mask_bytes=0,
             ceiling=None):
    """Dominant stage kernel: algorithmic bytes per launch (stage 1 reads U, writes U*: 18
    values per interior cell; stage 2 reads U*, U, writes U: 27; with a mask, its byte per
    interior cell in each stage) over its mean duration. `ceiling`: mix_ceiling's measurement
    for the same cells, added as the dominant stage's streaming bandwidth and the fraction of
    it achieved."""
    per_cell = {0: 18 * elem_bytes + mask_bytes, 1: 27 * elem_bytes + mask_bytes}
    avg = [kernels[i][0] / max(1, kernels[i][1]) for i in (0, 1)]  # ms
    dom = 0 if avg[0] > avg[1] else 1
    bytes_launch = per_cell[dom] * cells
    achieved = bytes_launch / (avg[dom] * 1e-3) / 1e9
    step_gbs = (45 * elem_bytes + 2 * mask_bytes) * cells / (ms_per_step * 1e-3) / 1e9
    out = {
        "bound": "hbm", "kernel": f"{name}<{dom + 1}>", "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic(dom) if traffic else None,
        "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg[dom], 4),
        "stage_avg_ms": [round(a, 4) for a in avg],
        "step_achieved_gbs": round(step_gbs, 1), "step_frac": round(step_gbs / HBM_PEAK_GBS, 4),
    }
    if ceiling:
        out["stream_mix_gbs"] = ceiling[dom + 1]
        out["frac_of_stream_mix"] = round(achieved / ceiling[dom + 1], 4)
    return out


# ---- configured single runs (Orszag-Tang 256^2 to t = 0.5; rotor 1024^2 non-ideal + walls) ---

def config_leg(m, U0, n, precision, p, bc=(0, 0), mask=None, steps=None):
    """One configured solve on the single-run solver, on the graph path between stream
    synchronisations with the state resident on the device (uploaded before the timed region):
    `steps` steps, or to p.t_end. Before it, an untimed pass (autotuning of the shape, graph
    capture; for a solve to t_end it also gives the step count) and the same step count with
    events around every stage kernel. Returns the device solver too, holding the final state."""
    from mhd_gpu_solver_amd import _lib
    grid = _lib.Grid(n, n, 2, 0.0, 1.0, 0.0, 1.0)
    dev = m.DeviceSolver(grid, _lib.PRECISIONS[precision])
    dev.configure(bc[0], bc[1], mask, True)

    def start():
        dev.upload(U0)
        dev.fill_boundaries(p.gamma)

    start()
    count = steps if steps else dev.solve(p).steps   # warm-up (a whole solve to t_end)
    if steps:
        dev.advance_steps(p, min(steps, 8))
    start()
    dev.set_kernel_timing(True)
    dev.advance_steps(p, count)
    k = dev.kernel_times()
    dev.set_kernel_timing(False)
    start()
    t0 = time.perf_counter()
    r = dev.advance_steps(p, steps) if steps else dev.solve(p)
    secs = time.perf_counter() - t0
    cells = n * n
    return dict(mcups=cells * r.steps / secs / 1e6, ms_per_step=secs * 1e3 / r.steps,
                steps=r.steps, t=r.t, kernels=k, cells=cells, seg_rows=dev.seg_rows(), dev=dev)


def configs_leg(m, workloads):
    """Orszag-Tang 256^2 (B0 = 1/sqrt(4 pi), minmod, periodic) to t = 0.5, and the rotor 1024^2
    with Ohmic, Hall and ambipolar terms, outflow and 8 seeded solid circles as reflective walls,
    200 steps: fp64 and FP32, the per-field relative L1 of FP32 against fp64 on the device."""
    out = {}
    p = params(m)
    p.t_end = 0.5
    U0 = m.init_ic("orszag_tang_b0", 256, 256, gamma=5.0 / 3.0, b0=B0_STD)
    legs = {prec: config_leg(m, U0, 256, prec, p) for prec in ("fp64", "fp32")}
    out["orszag_tang_256_t0.5"] = _config_summary(m, legs, mask_bytes=0)
    n = 1024
    spec = workloads.rotor_circles(n)
    p = params(m)
    p.nonideal = spec.nonideal
    legs = {prec: config_leg(m, spec.state(), n, prec, p, bc=spec.bc, mask=spec.mask, steps=200)
            for prec in ("fp64", "fp32")}
    out["rotor_1024_nonideal_8_circles_200_steps"] = _config_summary(m, legs, mask_bytes=1)
    return out


def _config_summary(m, legs, mask_bytes):
    res = {}
    for prec, leg in legs.items():
        eb = 8 if prec == "fp64" else 4
        res[prec] = {"value": round(leg["mcups"], 2), "ms_per_step": round(leg["ms_per_step"], 4),
                     "steps": leg["steps"], "t": leg["t"], "seg_rows": leg["seg_rows"],
                     "roofline": roofline(leg["kernels"], leg["cells"], leg["ms_per_step"], eb,
                                          "k_stage", mask_bytes=mask_bytes)}
    l1 = legs["fp32"]["dev"].rel_l1(legs["fp64"]["dev"])
    res["fp32_rel_l1_vs_fp64"] = {f: float(f"{v:.3e}") for f, v in zip(m.FIELDS[:8], l1[:8])}
    return res


# ---- ensemble -------------------------------------------------------------------------------

# Untimed steps of the timed instances themselves, run just before their timed steps: the
# admission of the ensemble (host states into the arena) leaves the device idle for tens of
# milliseconds, and the clocks that drop meanwhile take about two replays to come back, which
# a run in service never sees. The timed region is still exactly `steps` steps of every
# instance (its steps LEAD_STEPS + 1 .. LEAD_STEPS + steps).
LEAD_STEPS = 8


def ensemble_batched(m, specs, states, precision, steps, warmup, seg_rows=0):
    """The whole ensemble in one batch engine: `warmup` steps on each launch path (eager with
    events, then the graph: tiling, capture), then `steps` steps timed with events around every
    stage kernel, then the same on the graph path, whose rate is the leg's value. Each timed
    region follows LEAD_STEPS untimed steps of the same instances (after their admission) and
    the planning of the admitted set (`prepare`)."""
    budget = sum(m.instance_bytes(precision, s.n, s.n) for s in specs) + (64 << 20)
    eng = m.BatchEngine(precision, limiter=m.SlopeLimiter.Minmod, hbm_budget=budget,
                        max_instances=len(specs))
    if seg_rows:
        eng.set_seg_rows(seg_rows)

    def admit(k):
        return [eng.add_steps(U, k, gamma=s.gamma, bc=s.bc) for s, U in zip(specs, states)]

    for timing in (True, False):
        ids = admit(warmup)
        eng.set_kernel_timing(timing)
        eng.advance()
        for i in ids:
            eng.remove(i)
    t0 = time.perf_counter()
    ids = admit(LEAD_STEPS + steps)                  # host fp64 states into the HBM arena
    admit_s = time.perf_counter() - t0
    before = {i: eng.totals(i) for i, s in zip(ids, specs) if s.bc == (0, 0)}
    eng.advance(max_replays=1)                       # LEAD_STEPS untimed steps
    eng.set_kernel_timing(True)
    eng.prepare()                                    # the admitted set's tile tables
    t0 = time.perf_counter()
    eng.advance()
    wall = time.perf_counter() - t0
    kernels = eng.kernel_times()
    eng.set_kernel_timing(False)
    stats = eng.stats()
    drift = [0.0, 0.0]
    for i, t_b in before.items():
        t_a = eng.totals(i)
        for k, c in enumerate((0, 4)):  # mass, energy of the periodic (closed) instances
            drift[k] = max(drift[k], abs(t_a[c] - t_b[c]) / abs(t_b[c]))
    results = [eng.download(i) for i in ids]
    for i in ids:
        eng.remove(i)
    ids = admit(LEAD_STEPS + steps)
    eng.advance(max_replays=1)                       # LEAD_STEPS untimed steps
    t0 = time.perf_counter()
    eng.prepare()                                    # the admitted set's tables and graph
    prepare_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    eng.advance()
    graph_s = time.perf_counter() - t0
    for i in ids:
        eng.remove(i)
    cells = sum(s.n * s.n for s in specs)
    return dict(mcups=cells * steps / graph_s / 1e6, ms_per_step=graph_s * 1e3 / steps,
                timed_mcups=cells * steps / wall / 1e6, kernels=kernels, cells=cells,
                seg_rows=stats["seg_rows"], blocks=stats["blocks"], results=results,
                admit_s=admit_s, prepare_s=prepare_s,
                drift={"mass": float(f"{drift[0]:.3e}"), "energy": float(f"{drift[1]:.3e}"),
                       "instances": len(before)})


def ensemble_plain(m, specs, states, steps):
    """Each instance alone on the plain fp64 single-run solver (the reference's one-solve-per-
    DeviceSolver model); only the stepping calls are timed (state resident before each). Each
    instance runs LEAD_STEPS + `steps` steps, as in the batched legs, so that the results
    compare; the rate counts all of them."""
    devs = {}
    total = 0.0
    results = []
    p = params(m)
    for s, U in zip(specs, states):
        key = (s.n, s.bc)
        p.gamma = s.gamma
        if key not in devs:
            devs[key] = make_device(m, s.n, s.n, "fp64", s.bc)
            devs[key].upload(U)
            devs[key].advance_steps(p, 8)  # untimed: autotune and graph capture for this shape
        dev = devs[key]
        dev.upload(U)
        t0 = time.perf_counter()
        dev.advance_steps(p, LEAD_STEPS + steps)
        total += time.perf_counter() - t0
        out = np.zeros_like(U)
        dev.download(out)
        results.append(out)
    cells = sum(s.n * s.n for s in specs)
    return dict(mcups=cells * (LEAD_STEPS + steps) / total / 1e6, seconds=total,
                results=results)


def backend_compare(m, sizes, t_end, precisions=("fp64", "fp32"), repeats=3):
    """The reference's backend-comparison speed table (tools/mhd_backend_compare.cpp:128-153,
    509-531) on the hip backend: its Orszag-Tang (B0 = 1), MC limiter, to t_end; MHDSolver.run
    per size and precision, host<->device transfers of the state included, after a warm-up
    solve of the same shape (run(min(t_end, 1e-4)) on a second solver). The median of
    `repeats` runs, as the reference's median_rate takes (:290-300): runs of a few ms vary."""
    out = {}
    for prec in precisions:
        rows = []
        for n in sizes:
            def make():
                s = m.MHDSolver(n, n)
                s.set_limiter_mc()
                s.initialize("orszag_tang")
                s.set_backend("hip", prec)
                return s
            make().run(min(t_end, 1e-4))
            times = []
            for _ in range(repeats):
                s = make()
                t0 = time.perf_counter()
                r = s.run(t_end)
                times.append(time.perf_counter() - t0)
            secs = sorted(times)[len(times) // 2]
            rows.append({"n": n, "steps": r.steps, "seconds": round(secs, 4),
                         "mcups": round(n * n * r.steps / secs / 1e6, 1)})
        out[prec] = rows
    return out


def rel_l1_by_field(a, b):
    i = (slice(2, -2), slice(2, -2))
    out = []
    for c in range(8):
        den = np.abs(b[c][i]).sum()
        out.append(np.abs(a[c][i] - b[c][i]).sum() / den if den > 0 else 0.0)
    return out


# ---- CPU baseline -------------------------------------------------------------------------

def cpu_leg(specs, states, steps, budget_s=12.0):
    """The reference's CPU solver (Release flags) on this host, instance by instance over a
    bounded prefix of the ensemble, each for `steps` steps."""
    sys.path.insert(0, str(ROOT / "tests"))
    import ctypes as C

    from checkers import MINMOD, Reference, fast_reference_path, make_params

    path = fast_reference_path()
    if path is None:
        return None
    ref = Reference(path)
    threads = min(16, os.cpu_count() or 1)
    ref.lib.ref_set_threads(threads)
    done_cells = 0
    secs = 0.0
    count = 0
    for s, U0 in zip(specs, states):
        p = make_params(limiter=MINMOD, cfl=0.4, gamma=s.gamma, glm_alpha=0.1)
        U = np.ascontiguousarray(U0).copy()
        secs += ref.lib.ref_time_steps(s.n, s.n, 2, 0.0, 1.0, 0.0, 1.0, s.bc[0], s.bc[1],
                                       C.byref(p), steps, U)
        done_cells += s.n * s.n * steps
        count += 1
        if secs > budget_s:
            break
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(done_cells / secs / 1e6, 2), "unit": "Mcell-updates/s", "cores": threads,
        "kind": "reference",
        "host": {"cpu_model": model, "nproc": os.cpu_count(),
                 "isa": ref.lib.ref_simd_backend().decode()},
        "sample": f"first {count} of {len(specs)} ensemble instances, {steps} SSP-RK2 steps each, "
                  f"fp64, {secs:.1f} s, reference CPU solver ({path.name}, "
                  f"{ref.lib.ref_simd_backend().decode()} SIMD, {threads} threads)",
    }


def copy_probe(nbytes, repeats=3):
    """Plain device-to-device copies of known size (hipMemcpy's copy kernel): the scale for the
    FETCH_SIZE / WRITE_SIZE counters of a profiling run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed in copy probe")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    for _ in range(repeats):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed in copy probe")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {repeats} x {nbytes} B device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed steps (multiple of 8)")
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--single-n", type=int, default=4096)
    ap.add_argument("--single-steps", type=int, default=100)
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-slab", action="store_true",
                    help="skip the one-GPU slab of the single run (a profile's last launches "
                         "of each single-run kernel are then the full grid's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["ensemble", "single", "compare", "serve", "configs"],
                    default=None)
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the reference's backend-comparison table (OT, MC, t_end 0.02)")
    ap.add_argument("--serve-requests", type=int, default=128)
    ap.add_argument("--batch-rows", type=int, default=-1,
                    help="batch rows per wave: -1 timed on dry steps (default: the "
                         "instance set is static), 0 load model, >0 fixed")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the Orszag-Tang 256^2 and rotor 1024^2 non-ideal + walls runs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also run plain device copies of BYTES (counter scale for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the single-GPU throughput is the reported value)", file=sys.stderr)
        sys.exit(2)

    import mhd_gpu_solver_amd as m
    from mhd_gpu_solver_amd import workloads

    if not m.hip_available():
        raise SystemExit("no gfx950 device visible to the hip backend")
    if args.copy_probe:
        copy_probe(args.copy_probe)
    out = {"metric": METRIC, "unit": "Mcell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic"}

    if args.only in (None, "ensemble"):
        specs = workloads.ensemble(args.instances, args.seed)
        log(f"ensemble: {len(specs)} instances, {sum(s.n * s.n for s in specs)} cells")
        states = [s.state() for s in specs]
        log(f"batched fp32: {args.warmup} warmup + {args.steps} timed steps")
        b32 = ensemble_batched(m, specs, states, "fp32", args.steps, args.warmup,
                               args.batch_rows)
        log(f"batched fp32: {b32['mcups']:.1f} Mcell-updates/s ({b32['ms_per_step']:.3f} ms/step,"
            f" with kernel events {b32['timed_mcups']:.1f})")
        log("batched fp64")
        b64 = ensemble_batched(m, specs, states, "fp64", args.steps, args.warmup,
                               args.batch_rows)
        log(f"batched fp64: {b64['mcups']:.1f} Mcell-updates/s")
        log("batched fp64_fast")
        b64f = ensemble_batched(m, specs, states, "fp64_fast", args.steps, args.warmup,
                                args.batch_rows)
        log(f"batched fp64_fast: {b64f['mcups']:.1f} Mcell-updates/s")
        fast_l1 = max(max(rel_l1_by_field(a, b)) for a, b in zip(b64f["results"], b64["results"]))
        plain = None
        l1 = None
        if not args.no_plain:
            log("plain fp64, one instance at a time")
            plain = ensemble_plain(m, specs, states, args.steps)
            log(f"plain fp64: {plain['mcups']:.1f} Mcell-updates/s")
            worst = np.zeros(8)
            for a, b in zip(b32["results"], plain["results"]):
                worst = np.maximum(worst, rel_l1_by_field(a, b))
            l1 = {f: float(f"{v:.3e}") for f, v in zip(m.FIELDS[:8], worst)}
            # the batched fp64 engine reproduces the single-run fp64 kernel bit for bit
            same = all(np.array_equal(a, b) for a, b in zip(b64["results"], plain["results"]))
        ceil64 = mix_ceiling(m, "fp64", b64["cells"])
        out.update({
            "value": round(b32["mcups"], 2),
            "ms_per_step": round(b32["ms_per_step"], 4),
            "dtype": "f32",
            "config": {
                "workload": f"ensemble_{len(specs)}_mixed_batched_fp32",
                "instances": len(specs), "sizes": [64, 96, 128, 192, 256],
                "cells": b32["cells"], "ics": "orszag_tang_b0 (periodic) / rotor_b0 (outflow)",
                "gamma": "U[1.4,5/3]", "field_scale": "U[0.5,2]", "seed": args.seed,
                "limiter": "minmod", "riemann": "hll", "integrator": "ssp_rk2", "cfl": 0.4,
                "precision": "fp32 bulk, fp64 pressure recovery",
            },
            "timed_path_mcups": round(b32["timed_mcups"], 2),
            "roofline": roofline(b32["kernels"], b32["cells"], b32["ms_per_step"], 4,
                                 "k_batch_stage",
                                 traffic=lambda dom: measured_traffic(
                                     f"k_batch_stage<{dom + 1}", 4, f"ensemble_{len(specs)}",
                                     "k_batch_stage"),
                                 ceiling=mix_ceiling(m, "fp32", b32["cells"])),
            "seg_rows": b32["seg_rows"], "persistent_blocks": b32["blocks"],
            "row_choice": {-1: "timed", 0: "load model"}.get(args.batch_rows, "fixed"),
            "drift_fp32_periodic_max": b32["drift"],
            "admission_s": round(b32["admit_s"], 4),
            "planning_s": round(b32["prepare_s"], 4),
            "batched_fp64_fast": {"value": round(b64f["mcups"], 2),
                                  "ms_per_step": round(b64f["ms_per_step"], 4),
                                  "seg_rows": b64f["seg_rows"],
                                  "rel_l1_max_vs_fp64": float(f"{fast_l1:.3e}"),
                                  "roofline": roofline(
                                      b64f["kernels"], b64f["cells"], b64f["ms_per_step"], 8,
                                      "k_batch_stage",
                                      traffic=lambda dom: measured_traffic(
                                          f"k_batch_stage<{dom + 1}", 8,
                                          f"ensemble_{len(specs)}", "k_batch_stage", fast=1),
                                      ceiling=ceil64)},
            "batched_fp64": {"value": round(b64["mcups"], 2),
                             "drift_periodic_max": b64["drift"],
                             "ms_per_step": round(b64["ms_per_step"], 4),
                             "seg_rows": b64["seg_rows"],
                             "roofline": roofline(
                                 b64["kernels"], b64["cells"], b64["ms_per_step"], 8,
                                 "k_batch_stage",
                                 traffic=lambda dom: measured_traffic(
                                     f"k_batch_stage<{dom + 1}", 8, f"ensemble_{len(specs)}",
                                     "k_batch_stage"),
                                 ceiling=ceil64)},
        })
        if plain:
            out["plain_fp64"] = {"value": round(plain["mcups"], 2),
                                 "seconds": round(plain["seconds"], 3),
                                 "model": "single-run fp64 solver, one instance after another"}
            out["batched_fp32_vs_plain_fp64"] = round(b32["mcups"] / plain["mcups"], 2)
            out["fp32_rel_l1_max_vs_fp64"] = l1
            out["fp32_l1_within_1e-4"] = bool(max(l1.values()) <= 1e-4)
            out["batched_fp64_equals_single_run_fp64"] = bool(same)
        if not args.no_cpu:
            log("cpu baseline (reference build, bounded ensemble sample)")
            out["cpu_baseline"] = cpu_leg(specs, states, args.steps)
            if out["cpu_baseline"]:
                log(f"cpu: {out['cpu_baseline']['value']} Mcell-updates/s")
        else:
            out["cpu_baseline"] = None

    if not args.no_single and args.only in (None, "single"):
        n = args.single_n
        log(f"single run {n}x{n}")
        U0 = m.init_ic("orszag_tang_b0", n, n, gamma=5.0 / 3.0, b0=B0_STD)
        legs = {}
        ceilings = {eb: mix_ceiling(m, "fp64" if eb == 8 else "fp32", n * n) for eb in (8, 4)}
        for prec, eb in (("fp64", 8), ("fp32", 4), ("fp64_fast", 8)):
            leg = single_leg(m, n, prec, args.single_steps, args.warmup, U0)
            log(f"single {prec}: {leg['mcups']:.1f} Mcell-updates/s ({leg['ms_per_step']:.3f} ms/step)")
            legs[prec] = {
                "value": round(leg["mcups"], 2), "ms_per_step": round(leg["ms_per_step"], 4),
                "timed_path_mcups": round(leg["timed_mcups"], 2), "seg_rows": leg["seg_rows"],
                "drift": leg["drift"], "pcie_inclusive": leg["pcie"],
                "roofline": roofline(leg["kernels"], leg["cells"], leg["ms_per_step"], eb,
                                     "k_stage",
                                     traffic=lambda dom, eb=eb, prec=prec: measured_traffic(
                                         f"k_stage<{dom + 1}", eb, f"orszag_tang_{n}",
                                         fast=int(prec == "fp64_fast")),
                                     ceiling=ceilings[eb]),
            }
        out["single_run"] = {"workload": f"orszag_tang_{n}x{n}_single_run", **legs}
        # One GPU's share of the same run split into 8 y-slabs: a n x n/8 slab of the state
        # (its ghost rows the neighbouring slabs' rows at t = 0), periodic in x; outflow in y
        # stands in for the halo rows a multi-GPU run would receive before each stage.
        slabs = 8
        if n % slabs == 0 and not args.no_slab:
            h = n // slabs
            Us = np.ascontiguousarray(U0[:, :h + 4, :])
            slab = {"shape": f"{n}x{h}", "of": slabs, "bc": "periodic x, outflow y"}
            for prec, eb in (("fp64", 8), ("fp32", 4)):
                leg = single_leg(m, n, prec, args.single_steps, args.warmup, Us, ny=h,
                                 bc=(0, 2), y_max=h / n)
                log(f"slab {n}x{h} {prec}: {leg['mcups']:.1f} Mcell-updates/s")
                slab[prec] = {
                    "value": round(leg["mcups"], 2), "ms_per_step": round(leg["ms_per_step"], 4),
                    "seg_rows": leg["seg_rows"],
                    "roofline": roofline(leg["kernels"], leg["cells"], leg["ms_per_step"], eb,
                                         "k_stage"),
                }
            out["single_run"]["slab"] = slab
        if args.only == "single":
            out.update({"value": legs["fp64"]["value"], "ms_per_step": legs["fp64"]["ms_per_step"],
                        "dtype": "f64", "config": {"workload": f"orszag_tang_{n}x{n}_single_run"},
                        "roofline": legs["fp64"]["roofline"], "cpu_baseline": None})
    if not args.no_configs and args.only == "configs":
        log("configured runs: Orszag-Tang 256^2 to t = 0.5; rotor 1024^2 non-ideal + walls")
        cfg = configs_leg(m, workloads)
        for name, res in cfg.items():
            log(f"{name}: fp64 {res['fp64']['value']} fp32 {res['fp32']['value']} Mcell-updates/s, "
                f"{res['fp64']['steps']} steps")
        out["configs"] = cfg
        if args.only == "configs":
            top = cfg["rotor_1024_nonideal_8_circles_200_steps"]["fp32"]
            out.update({"value": top["value"], "dtype": "f32", "ms_per_step": top["ms_per_step"],
                        "config": {"workload": "rotor_1024_nonideal_8_circles_200_steps"},
                        "roofline": top["roofline"], "cpu_baseline": None})
    if not args.no_configs and args.only is None:
        # the configured runs in a child process of their own: a failure there (reported in
        # the line) cannot take the line with it
        import subprocess
        log("configured runs (child process)")
        try:
            child = subprocess.run([sys.executable, str(pathlib.Path(__file__).resolve()),
                                    "--only", "configs"], capture_output=True, text=True,
                                   timeout=600)
            sys.stderr.write(child.stderr)
            out["configs"] = json.loads(child.stdout.strip().splitlines()[-1])["configs"]
        except subprocess.TimeoutExpired:
            out["configs"] = {"error": "timed out"}
        except (IndexError, KeyError, ValueError):
            out["configs"] = {"error": f"exit status {child.returncode}"}
    if not args.no_compare and args.only in (None, "compare"):
        log("backend comparison: Orszag-Tang (B0 = 1), MC, t_end 0.02, transfers included")
        cmp_ = backend_compare(m, (128, 256, 512, 1024), 0.02)
        for prec, rows in cmp_.items():
            log(f"compare {prec}: " + ", ".join(f"{r['n']}^2 {r['mcups']}" for r in rows))
        out["backend_compare"] = {
            "workload": "orszag_tang (reference IC, B0 = 1), MC, t_end 0.02, MHDSolver.run "
                        "incl. host<->device transfers",
            **cmp_}
        if args.only == "compare":
            top = cmp_["fp32"][-1]
            out.update({"value": top["mcups"], "dtype": "f32", "ms_per_step": None,
                        "config": {"workload": f"backend_compare_orszag_tang_{top['n']}"},
                        "roofline": None, "cpu_baseline": None})
    if not args.no_serve and args.only in (None, "serve"):
        from mhd_gpu_solver_amd.serve import Server
        reqs = workloads.request_stream(args.serve_requests, rate=200.0, seed=7)
        log(f"serving: {len(reqs)} requests, Poisson 200/s, N~U{{64..1024}}, t_end~U[0.05,0.5]")
        runs = {}
        for policy in ("srpt", "all"):
            srv = Server("fp32", hbm_budget=32 << 30, max_instances=1024, policy=policy)
            summ = srv.run(reqs, realtime=True).summary()
            del srv
            log(f"serving ({policy}): p50 {summ['p50_ms']} ms, p99 {summ['p99_ms']} ms, "
                f"{summ['mcups']} Mcell-updates/s, max concurrent {summ['max_concurrent']}")
            runs[policy] = {"policy": policy, **summ}
        summ = runs["srpt"]
        out["serving"] = {"workload": "request_stream_poisson_200", "precision": "fp32",
                          "limiter": "mc", "hbm_budget_bytes": 32 << 30,
                          "work_cap_cells": Server.DEFAULT_WORK_CAP, **summ,
                          "processor_sharing": runs["all"]}
        if args.only == "serve":
            out.update({"value": summ["mcups"], "dtype": "f32", "config": {
                "workload": "request_stream_poisson_200"}, "roofline": None,
                "cpu_baseline": None, "ms_per_step": None})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
