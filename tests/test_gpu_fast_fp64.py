"""The fast fp64 build (precision "fp64_fast": contracted multiply-adds, Newton-refined
hardware reciprocal and square root in the stepping kernels) against the oracle: within 1e-12
relative L1 per conserved field after a fixed step count at the same dt sequence (the oracle driven by the GPU's logged dt: orc_solve_dt; the GLM scalar psi,
which stays at round-off size, is bounded against |B| instead). Also the
reference's fp64 GPU gate (tools/mhd_backend_compare.cpp:242-250: equal step counts, relative
L1 <= 1e-11, L-inf <= 1e-9, drift <= 1e-11) and the batched engine, masks and non-ideal terms
included."""
import math

import numpy as np
import pytest

from checkers import FIELDS, MC, MINMOD, OUTFLOW, make_grid, make_params

pytestmark = pytest.mark.gpu

BAR = 1e-12


def rel_l1(a, b):
    out = {}
    for c, name in enumerate(FIELDS):
        den = np.abs(b[c]).sum()
        num = np.abs(a[c] - b[c]).sum()
        out[name] = num / den if den > 0 else num
    return out


@pytest.mark.slow
def test_orszag_tang_4096_50_steps_at_the_same_dt(hip, oracle):
    n, b0 = 4096, 1.0 / math.sqrt(4 * math.pi)
    s = hip.MHDSolver(n, n)
    s.set_backend("hip", "fp64_fast")
    s.set_limiter_minmod()
    s.initialize("orszag_tang_b0", b0=b0)
    U0 = s.state.copy()
    dev = s._dev()
    assert dev.precision() == "fp64_fast"
    dev.set_dt_log(64)
    r = s.advance_steps(50)
    dts = dev.dt_log(64)
    assert r.steps == 50 and len(dts) == 50
    p = make_params(t_end=1e300, cfl=0.4, limiter=MINMOD, max_steps=50)
    Uo, info = oracle.solve_dt(make_grid(n, n), U0, p, dts)
    l1 = rel_l1(s.state, Uo)
    # psi, the GLM cleaning scalar, starts at 0 and stays at round-off size (the initial field
    # is divergence-free), so its relative L1 measures noise: its difference is bounded
    # against the field instead
    psi_abs = np.abs(s.state[8] - Uo[8]).sum() / np.abs(Uo[5:7]).sum()
    print("fp64_fast OT 4096^2 50 steps, rel L1 at the same dt:",
          " ".join(f"{k} {v:.2e}" for k, v in l1.items()), f"| psi / |B| {psi_abs:.2e}")
    assert max(l1[f] for f in FIELDS[:8]) <= BAR, l1
    assert psi_abs <= BAR
    # its own dt sequence stays within rounding of the exact one
    _, ex = oracle.solve(make_grid(n, n), U0, p, allow_max_steps=True, dt_cap=64)
    assert np.max(np.abs(dts / ex["dt"] - 1.0)) <= 1e-12


@pytest.mark.parametrize("nonideal", [False, True], ids=["ideal", "ohmic_hall"])
def test_fp64_gate_orszag_tang_128(hip, oracle, nonideal):
    from test_gpu_gate import compare_states, setup_orszag
    n, t_acc = 128, 0.25
    s = hip.MHDSolver(n, n)
    setup_orszag(s, nonideal)
    s.set_backend("hip", "fp64_fast")
    U0 = s.state.copy()
    r = s.run(t_acc)
    p = make_params(t_end=t_acc, limiter=MC, eta_ohm=1e-3 if nonideal else None,
                    eta_hall=5e-4 if nonideal else None)
    Uo, info = oracle.solve(make_grid(n, n), U0, p)
    l1, linf = compare_states(s.state, Uo, 5.0 / 3.0)
    i = (slice(2, -2), slice(2, -2))
    dm = abs(s.state[0][i].sum() - U0[0][i].sum()) / U0[0][i].sum()
    de = abs(s.state[4][i].sum() - U0[4][i].sum()) / U0[4][i].sum()
    print(f"gate fp64_fast OT128 {'ohmic+hall' if nonideal else 'ideal'}: steps {r.steps} vs "
          f"{info['steps']}, L1 {l1:.2e}, Linf {linf:.2e}, drift {dm:.1e} / {de:.1e}")
    assert r.steps == info["steps"]
    assert l1 <= 1e-11 and linf <= 1e-9 and max(dm, de) <= 1e-11


def test_rotor_nonideal_mask_fast_at_the_same_dt(hip, oracle):
    """Ohmic + Hall + ambipolar, outflow and solid circles with reflective walls, 30 steps."""
    from mhd_gpu_solver_amd import workloads
    n = 256
    mask = workloads.circles_mask(n, 8, seed=11)
    s = hip.MHDSolver(n, n)
    s.set_backend("hip", "fp64_fast")
    s.set_outflow_bc()
    s.set_limiter_minmod()
    s.enable_ohmic(1e-3)
    s.enable_hall(1e-2)
    s.enable_ambipolar(1e-3)
    s.initialize("rotor")
    s.set_mask(mask)
    U0 = s.state.copy()
    dev = s._dev()
    dev.set_dt_log(64)
    s.advance_steps(30)
    dts = dev.dt_log(64)
    p = make_params(t_end=1e300, limiter=MINMOD, eta_ohm=1e-3, eta_hall=1e-2,
                    eta_ambipolar=1e-3)
    Uo, _ = oracle.solve_dt(make_grid(n, n), U0, p, dts, bc=(OUTFLOW, OUTFLOW), mask=mask)
    l1 = rel_l1(s.state, Uo)
    assert max(l1[f] for f in FIELDS[:8]) <= BAR, l1


def test_batched_fast_ensemble_sample_within_bar_of_exact(hip):
    """fp64_fast batch against the exact fp64 batch on a sample of the mixed-size ensemble
    (8 steps; their dt sequences agree to rounding)."""
    from mhd_gpu_solver_amd import workloads
    specs = workloads.ensemble(512, 2024)[:48]
    states = [sp.state() for sp in specs]
    engines = {}
    for prec in ("fp64", "fp64_fast"):
        eng = hip.BatchEngine(prec, limiter=hip.SlopeLimiter.Minmod, hbm_budget=1 << 31,
                              max_instances=64)
        ids = [eng.add_steps(U, 8, gamma=sp.gamma, bc=sp.bc) for sp, U in zip(specs, states)]
        eng.advance()
        engines[prec] = (eng, ids)
    worst = 0.0
    (ef, idf), (ee, ide) = engines["fp64_fast"], engines["fp64"]
    for a, b in zip(idf, ide):
        ra, rb = ef.result(a)[0], ee.result(b)[0]
        assert ra.steps == rb.steps == 8 and abs(ra.t / rb.t - 1.0) <= 1e-12
        worst = max(worst, float(ef.rel_l1(a, ee, b)[:8].max()))
    print(f"fp64_fast batch vs fp64 batch, worst rel L1: {worst:.2e}")
    assert worst <= BAR
