"""Benchmark of the hip backend's hot path: SSP-RK2 steps of the GLM-MHD MUSCL/HLL stencil.

Default workload (one GPU): Orszag-Tang vortex at 4096 x 4096 on [0,1]^2, periodic, gamma 5/3,
B0 = 1/sqrt(4 pi), minmod, HLL, CFL 0.4, glm_alpha 0.1. A "step" is one full SSP-RK2 step of
the whole grid (two fused stage kernels, CFL reduction fused into stage 2). The state is
resident in HBM before the timed region starts; upload/download are outside it.

value: fp64 Mcell-updates/s = nx * ny * steps / seconds / 1e6 (the reference's definition,
tools/mhd_backend_compare.cpp:295,527). The FP32-bulk engine on the same workload is reported
beside it. roofline: the dominant stage kernel's algorithmic bytes per launch over its average
duration, from HIP events recorded on the solver's stream around each launch in the timed
region. cpu_baseline: the reference's own CPU solver (Release flags) timed on this host on a
bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx N] [--no-cpu]
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


def make_device(m, nx, ny, precision):
    from mhd_gpu_solver_amd import _lib
    grid = _lib.Grid(nx, ny, 2, 0.0, 1.0, 0.0, 1.0)
    dev = m.DeviceSolver(grid, _lib.FP64 if precision == "fp64" else _lib.FP32)
    dev.configure(0, 0, None, True)
    return dev


def initial_state(m, nx, ny):
    U = m.init_ic("orszag_tang_b0", nx, ny, gamma=5.0 / 3.0, b0=B0_STD)
    return U


def params(m):
    p = m.SolveParams()
    p.limiter = m.SlopeLimiter.Minmod
    p.cfl = 0.4
    p.gamma = 5.0 / 3.0
    p.glm_alpha = 0.1
    return p


def gpu_leg(m, nx, ny, precision, steps, warmup, U0):
    """Times `steps` steps after `warmup`; returns (mcups, ms_per_step, kernel stats, graph mcups)."""
    dev = make_device(m, nx, ny, precision)
    dev.upload(U0)
    dev.fill_boundaries(5.0 / 3.0)
    p = params(m)
    if warmup:
        dev.advance_steps(p, warmup)
    # production path (graph replays), wall clock around a synchronising call
    t0 = time.perf_counter()
    dev.advance_steps(p, steps)
    graph_s = time.perf_counter() - t0
    # timed region with an event pair around every stage kernel
    dev.set_kernel_timing(True)
    t0 = time.perf_counter()
    dev.advance_steps(p, steps)
    wall = time.perf_counter() - t0
    k = dev.kernel_times()
    rows = dev.seg_rows()
    dev.set_kernel_timing(False)
    U = np.zeros_like(U0)
    dev.download(U)
    if not np.isfinite(U).all():
        raise RuntimeError(f"{precision}: non-finite state after the timed steps")
    cells = nx * ny
    return dict(mcups=cells * steps / wall / 1e6, ms_per_step=wall * 1e3 / steps,
                graph_mcups=cells * steps / graph_s / 1e6, kernels=k, wall=wall, seg_rows=rows)


def measured_traffic(kernel_prefix, nx, ny, elem_bytes):
    """HBM bytes per launch of a stage kernel from the committed FETCH_SIZE/WRITE_SIZE summary
    (profiles/traffic_k_stage_orszag_tang_<n>.json), or None when none exists for this size."""
    path = ROOT / "profiles" / f"traffic_k_stage_orszag_tang_{nx}.json"
    if nx != ny or not path.exists():
        return None
    data = json.loads(path.read_text())
    want = "double" if elem_bytes == 8 else "float"
    for name, v in data.get("kernels", {}).items():
        if name.startswith(kernel_prefix) and want in name:
            return v["traffic_bytes"]
    return None


def roofline(leg, nx, ny, elem_bytes):
    cells = nx * ny
    # algorithmic bytes per launch: stage 1 reads U, writes U*; stage 2 reads U*, U, writes U
    per_cell = {0: 18 * elem_bytes, 1: 27 * elem_bytes}
    avg = [leg["kernels"][i][0] / max(1, leg["kernels"][i][1]) for i in (0, 1)]  # ms
    dom = 0 if avg[0] > avg[1] else 1
    bytes_launch = per_cell[dom] * cells
    achieved = bytes_launch / (avg[dom] * 1e-3) / 1e9
    step_gbs = 45 * elem_bytes * cells / (leg["ms_per_step"] * 1e-3) / 1e9
    return {
        "bound": "hbm",
        "kernel": f"k_stage<{dom + 1}>",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(f"k_stage<{dom + 1}", nx, ny, elem_bytes),
        "bytes_per_launch": bytes_launch,
        "avg_launch_ms": round(avg[dom], 4),
        "seg_rows": leg["seg_rows"],
        "stage_avg_ms": [round(a, 4) for a in avg],
        "step_achieved_gbs": round(step_gbs, 1),
        "step_frac": round(step_gbs / HBM_PEAK_GBS, 4),
    }


def cpu_leg(m, nx, ny, U0, budget_s=12.0):
    """Reference Release build on this host's cores, bounded sample of the same workload."""
    sys.path.insert(0, str(ROOT / "tests"))
    import ctypes as C

    from checkers import MINMOD, Reference, fast_reference_path, make_params

    path = fast_reference_path()
    if path is None:
        return None
    ref = Reference(path)
    threads = min(16, os.cpu_count() or 1)
    ref.lib.ref_set_threads(threads)
    p = make_params(limiter=MINMOD, cfl=0.4, gamma=5.0 / 3.0, glm_alpha=0.1)
    U = np.ascontiguousarray(U0).copy()

    def run(k):
        return ref.lib.ref_time_steps(nx, ny, 2, 0.0, 1.0, 0.0, 1.0, 0, 0, C.byref(p), k, U)

    run(1)  # thread pool start-up, first touch
    t2 = run(2)
    n = int(max(2, min(200, budget_s / max(t2 / 2, 1e-6))))
    secs = run(n)
    return {
        "value": round(nx * ny * n / secs / 1e6, 2),
        "unit": "Mcell-updates/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"orszag_tang {nx}x{ny} fp64 minmod, {n} SSP-RK2 steps in {secs:.1f} s, "
                  f"reference CPU solver ({path.name}, {ref.lib.ref_simd_backend().decode()} "
                  f"SIMD, {threads} threads)",
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
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="also run plain device copies of BYTES (counter scale for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the single-GPU throughput is the reported value)", file=sys.stderr)
        sys.exit(2)

    import mhd_gpu_solver_amd as m

    if not m.hip_available():
        raise SystemExit("no gfx950 device visible to the hip backend")
    if args.copy_probe:
        copy_probe(args.copy_probe)
    nx = ny = args.nx
    log(f"building initial state {nx}x{ny}")
    U0 = initial_state(m, nx, ny)
    log(f"fp64: {args.warmup} warmup + {args.steps} timed steps")
    f64 = gpu_leg(m, nx, ny, "fp64", args.steps, args.warmup, U0)
    log(f"fp64: {f64['mcups']:.1f} Mcell-updates/s ({f64['ms_per_step']:.3f} ms/step)")
    f32 = None
    if not args.no_fp32:
        log(f"fp32-bulk: {args.warmup} warmup + {args.steps} timed steps")
        f32 = gpu_leg(m, nx, ny, "fp32", args.steps, args.warmup, U0)
        log(f"fp32: {f32['mcups']:.1f} Mcell-updates/s ({f32['ms_per_step']:.3f} ms/step)")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (reference build, bounded sample)")
        # the CPU leg starts from the device-filled initial state
        dev = make_device(m, nx, ny, "fp64")
        dev.upload(U0)
        dev.fill_boundaries(5.0 / 3.0)
        Uf = np.zeros_like(U0)
        dev.download(Uf)
        del dev
        cpu = cpu_leg(m, nx, ny, Uf)
        if cpu:
            log(f"cpu: {cpu['value']} Mcell-updates/s ({cpu['sample']})")
    out = {
        "metric": METRIC,
        "value": round(f64["mcups"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(f64["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"orszag_tang_{nx}x{ny}_single_run",
            "nx": nx, "ny": ny, "ic": "orszag_tang_b0", "b0": "1/sqrt(4 pi)",
            "limiter": "minmod", "riemann": "hll", "integrator": "ssp_rk2", "cfl": 0.4,
            "bc": "periodic", "precision": "fp64",
        },
        "graph_path_mcups": round(f64["graph_mcups"], 2),
        "roofline": roofline(f64, nx, ny, 8),
        "cpu_baseline": cpu,
    }
    if f32:
        out["fp32_bulk"] = {
            "value": round(f32["mcups"], 2),
            "ms_per_step": round(f32["ms_per_step"], 4),
            "graph_path_mcups": round(f32["graph_mcups"], 2),
            "speedup_vs_fp64": round(f32["mcups"] / f64["mcups"], 3),
            "roofline": roofline(f32, nx, ny, 4),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
