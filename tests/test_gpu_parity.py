"""Parity of the hip backend against the CPU oracle (the reference's CPU fp64 path restated).

Mirrors the reference's GPU-backend tests (tests/test_gpu_backends.cpp:16-157), with tighter bars:
fp64 must reproduce the oracle's dt sequence bit for bit and every conserved field to
1e-12 relative L1 (in practice bit-exact); fp32-bulk must stay within 1e-4 relative L1 per field.
Each case starts the device and the oracle from the same host state.
"""
import pathlib

import numpy as np
import pytest

from checkers import FIELDS, make_grid, make_params

GOLD = pathlib.Path(__file__).parent / "golden"

pytestmark = pytest.mark.gpu

FP64_L1 = 1e-12   # per conserved field, relative L1, fp64 vs oracle
FP32_L1 = 1e-4    # per conserved field, relative L1, fp32-bulk vs oracle


def oracle_params(s, t_end, max_steps=None):
    p = s.params
    ni = p.nonideal
    return make_params(t_end=t_end, cfl=p.cfl, gamma=p.gamma, glm_alpha=p.glm_alpha,
                       limiter=int(p.limiter),
                       eta_ohm=ni.eta_ohm if ni.ohmic else None,
                       eta_hall=ni.eta_hall if ni.hall else None,
                       eta_ambipolar=ni.eta_ambipolar if ni.ambipolar else None,
                       max_steps=max_steps if max_steps is not None else p.max_steps)


def grid_of(s):
    return make_grid(s.nx, s.ny, s.x_min, s.x_max, s.y_min, s.y_max, s.ng)


def rel_l1(a, b):
    den = np.abs(b).sum()
    num = np.abs(a - b).sum()
    return num / den if den > 0 else num


def run_pair(hip, oracle, nx, ny, setup, t_end, precision="fp64", y_max=None, dt_cap=4096):
    s = hip.MHDSolver(nx, ny, 0.0, 1.0, 0.0, (1.0 / nx if ny == 1 else 1.0) if y_max is None
                      else y_max)
    s.set_backend("hip", precision)
    setup(s)
    U0 = s.state.copy()
    dev = s._dev()
    dev.set_dt_log(dt_cap)
    r = s.run(t_end)
    dts = dev.dt_log(dt_cap)
    mask = s._active_mask()
    Uo, info = oracle.solve(grid_of(s), U0, oracle_params(s, t_end), bc=tuple(int(b) for b in s._bc),
                            mask=mask, walls=s.walls_reflective, dt_cap=dt_cap)
    return s, r, dts, Uo, info


def assert_fp64_parity(s, r, dts, Uo, info):
    assert r.steps == info["steps"]
    assert r.t == info["t"]
    assert r.c_h == info["c_h"]
    np.testing.assert_array_equal(dts, info["dt"])
    for c, name in enumerate(FIELDS):
        err = rel_l1(s.state[c], Uo[c])
        assert err <= FP64_L1, f"{name}: rel L1 {err:.3e}"
    # in practice the fused kernels round exactly like the oracle
    assert np.array_equal(s.state, Uo), f"max |diff| {np.abs(s.state - Uo).max():.3e}"


def test_initialize_matches_oracle_bitwise(hip, oracle):
    s = hip.MHDSolver(48, 40)
    s.set_backend("hip", "fp64")
    s.initialize("orszag_tang")
    Uo = oracle.init_ic("orszag_tang", make_grid(48, 40))
    assert np.array_equal(s.state, Uo)


def test_orszag_tang_mc_fp64_bitexact(hip, oracle):
    def setup(s):
        s.set_limiter_mc()
        s.initialize("orszag_tang")
    assert_fp64_parity(*run_pair(hip, oracle, 48, 40, setup, 0.05))


def test_sod_minmod_outflow_strip_fp64_bitexact(hip, oracle):
    def setup(s):
        s.set_limiter_minmod()
        s.set_bc(hip.BoundaryConditionType.Outflow, hip.BoundaryConditionType.Periodic)
        s.initialize("sod")
    assert_fp64_parity(*run_pair(hip, oracle, 64, 1, setup, 0.05))


def test_nonideal_all_terms_fp64_bitexact(hip, oracle):
    def setup(s):
        s.initialize("orszag_tang")
        s.enable_ohmic(2e-3)
        s.enable_hall(1e-3)
        s.enable_ambipolar(5e-4)
    assert_fp64_parity(*run_pair(hip, oracle, 32, 32, setup, 0.03))


def test_solid_mask_fp64_bitexact(hip, oracle):
    def setup(s):
        s.initialize("orszag_tang")
        m = np.ones((40, 40), dtype=np.uint8)
        m[15:25, 12:20] = 0
        s.set_mask(m)
    assert_fp64_parity(*run_pair(hip, oracle, 40, 40, setup, 0.03))


def test_mask_touching_boundary_outflow_fp64_bitexact(hip, oracle):
    # solid cells on the domain edge: ghosts copy the solid's stale value before the wall fill
    def setup(s):
        s.set_outflow_bc()
        s.initialize("rotor")
        m = np.ones((36, 44), dtype=np.uint8)
        m[0:5, 0:3] = 0
        m[30:36, 40:44] = 0
        m[10, 20] = 0          # isolated single solid cell
        m[18:21, 5:8] = 0
        m[19, 6] = 0
        s.set_mask(m)
    assert_fp64_parity(*run_pair(hip, oracle, 44, 36, setup, 0.02))


def test_fp32_bulk_within_tolerance(hip, oracle):
    def setup(s):
        s.set_limiter_mc()
        s.initialize("orszag_tang")
    s, r, dts, Uo, info = run_pair(hip, oracle, 48, 40, setup, 0.05, precision="fp32")
    assert abs(r.steps - info["steps"]) <= 1
    for c, name in enumerate(FIELDS[:8]):
        err = rel_l1(s.state[c], Uo[c])
        assert err <= FP32_L1, f"{name}: rel L1 {err:.3e}"


@pytest.mark.parametrize("rows", [1, 3, 7, 64])
def test_results_independent_of_tiling(hip, oracle, rows):
    """Any rows-per-wave value (the autotuner's knob) gives the oracle's bits; odd sizes and
    strips that end inside the grid exercise the partial tiles."""
    nx, ny = 131, 37
    s = hip.MHDSolver(nx, ny)
    s.set_backend("hip", "fp64")
    s.set_limiter_mc()
    s.initialize("rotor")
    s.enable_ohmic(1e-3)
    s.enable_hall(2e-3)
    U0 = s.state.copy()
    dev = s._dev()
    dev.set_seg_rows(rows)
    r = s.run(0.01)
    assert dev.seg_rows() == rows
    Uo, info = oracle.solve(grid_of(s), U0, oracle_params(s, 0.01))
    assert r.steps == info["steps"]
    assert np.array_equal(s.state, Uo)


@pytest.mark.parametrize("precision", ["fp32", "fp64_fast"])
def test_contracted_builds_independent_of_tiling(hip, precision):
    """The FP32 and contracted fp64 single runs give the same bits for every rows-per-wave
    value: odd row segments march the other way, and both directions must compute every cell
    alike (their fused multiply-adds are fixed in the source, not left to the compiler)."""
    nx, ny = 131, 37
    first = None
    for rows in (37, 1, 2, 3, 5, 8):
        s = hip.MHDSolver(nx, ny)
        s.set_backend("hip", precision)
        s.set_limiter_mc()
        s.initialize("rotor")
        s.enable_ohmic(1e-3)
        s.enable_hall(2e-3)
        dev = s._dev()
        dev.set_seg_rows(rows)
        r = s.run(0.01)
        assert dev.seg_rows() == rows
        if first is None:
            first = (r.steps, s.state.copy())
            continue
        assert r.steps == first[0], rows
        assert np.array_equal(s.state, first[1]), rows
    # and the ideal MC kernels, on an Orszag-Tang grid with strips that end inside it
    first = None
    for rows in (64, 1, 3, 6):
        s = hip.MHDSolver(125, 64)
        s.set_backend("hip", precision)
        s.set_limiter_mc()
        s.initialize("orszag_tang")
        dev = s._dev()
        dev.set_seg_rows(rows)
        r = s.run(0.02)
        if first is None:
            first = (r.steps, s.state.copy())
            continue
        assert r.steps == first[0], rows
        assert np.array_equal(s.state, first[1]), rows


def _compare_golden(got, ref, rtol, atol):
    """tests/test_regression_golden.cpp:125-137: per field rel L2 <= rtol, Linf <= max(atol,
    10 rtol); steps recorded identically."""
    assert got.meta["steps"] == ref.meta["steps"]
    for name, b in ref.fields.items():
        a = got.fields[name]
        assert a.size == b.size, name
        rl2 = np.sqrt(np.sum((a - b) ** 2)) / max(np.sqrt(np.sum(b ** 2)), 1e-300)
        assert rl2 <= rtol, (name, rl2)
        assert np.abs(a - b).max() <= max(atol, rtol * 10.0), name


def test_orszag_tang_32_golden_regression_on_device(hip):
    """The reference's OrszagTang32MatchesGolden recipe (test_regression_golden.cpp:50-79,
    140-156) on the hip backend, fp64, snapshot written through the package's golden I/O."""
    from mhd_gpu_solver_amd import state_io
    s = hip.MHDSolver(32, 32)
    s.set_backend("hip", "fp64")
    s.set_cfl(0.4)
    s.set_gamma(5.0 / 3.0)
    s.set_glm_alpha(0.1)
    s.set_limiter_mc()
    s.set_periodic_bc()
    s.set_ideal()
    s.initialize("orszag_tang")
    s.run(0.05)
    got = state_io.snapshot(s, "orszag_tang_32", ["rho", "bx", "by", "energy"],
                            {"t_end": "0.05", "cfl": "0.4", "gamma": "1.6666666666666667",
                             "glm_alpha": "0.1", "limiter": "MC"})
    ref = state_io.read_golden(GOLD / "orszag_tang_32.golden")
    _compare_golden(got, ref, 1e-8, 1e-10)
    assert got.meta == ref.meta


def test_ohmic_sine_bz_32_golden_regression_on_device(hip):
    """OhmicSineBz32MatchesGolden (test_regression_golden.cpp:82-125, 158-171) on hip fp64."""
    from mhd_gpu_solver_amd import state_io
    s = hip.MHDSolver(32, 32)
    s.set_backend("hip", "fp64")
    s.set_cfl(0.3)
    s.set_gamma(5.0 / 3.0)
    s.set_glm_alpha(0.1)
    s.set_limiter_minmod()
    s.set_periodic_bc()
    s.enable_ohmic(0.1)
    x = (np.arange(32) + 0.5) / 32.0
    one = np.ones((32, 32))
    zero = np.zeros((32, 32))
    s.set_primitives(one, zero, zero, zero, one, zero, zero,
                     np.broadcast_to(np.sin(2.0 * np.pi * x), (32, 32)))
    s.run(0.1)
    got = state_io.snapshot(s, "ohmic_sine_bz_32", ["bz", "energy"])
    ref = state_io.read_golden(GOLD / "ohmic_sine_bz_32.golden")
    _compare_golden(got, ref, 5e-8, 1e-9)


def test_device_name_and_precision(hip):
    """DeviceSolver::device_name / precision (device_solver.hpp:57-62) through the C ABI."""
    for precision in ("fp64", "fp32"):
        s = hip.MHDSolver(16, 16)
        s.set_backend("hip", precision)
        name = s.backend_device()
        assert name.startswith("HIP: ") and "gfx950" in name and "  " not in name, name
        assert s.precision == precision and s._dev().precision() == precision


def test_concurrent_solvers_from_threads(hip):
    """Several solvers may run at once from Python threads (the reference releases the GIL,
    python/mhd_pybind.cpp:279-290): each result equals the same solve run alone."""
    import threading

    def make(precision, ic, n):
        s = hip.MHDSolver(n, n)
        s.set_backend("hip", precision)
        s.set_limiter_mc()
        s.initialize(ic)
        return s

    cases = [("fp64", "orszag_tang", 96), ("fp32", "rotor", 80), ("fp64", "rotor", 64),
             ("fp32", "orszag_tang", 112)]
    alone = []
    for prec, ic, n in cases:
        s = make(prec, ic, n)
        r = s.run(0.02)
        alone.append((r.steps, s.state.copy()))
    solvers = [make(*c) for c in cases]
    out = [None] * len(cases)
    errors = []

    def work(k):
        try:
            r = solvers[k].run(0.02)
            out[k] = (r.steps, solvers[k].state.copy())
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    assert not errors, errors
    for (st_a, U_a), res in zip(alone, out):
        assert res is not None and res[0] == st_a
        assert np.array_equal(res[1], U_a)
