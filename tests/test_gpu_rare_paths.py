"""The face kernel's rare branches on the GPU, against the oracle bit for bit (fp64):

* the positivity / finiteness fallback (include/mhd/kernels/muscl_hll.hpp:145-168), reached
  by a block of cells with p < 0 (tests/rare_states.py);
* the degenerate wave-speed span average (:196-205), reached by a cold, slow flow whose
  global signal speed c_h (and so every face's span sR - sL) stays below 1e-14.
Both branches sit behind wave-uniform ballots in the kernels (taken when any lane whose flux
is used needs them), on the single-run and the batched paths. The FP32 path is held to the
1e-4 relative L1 bar on the fallback state; the cold flow is fp64 only, since its speeds'
fourth powers (the fast-speed discriminant) lie below the fp32 range."""
import numpy as np
import pytest

from checkers import MC, MINMOD, make_grid, make_params
from rare_states import cold_slow_flow, negative_pressure_patch

pytestmark = pytest.mark.gpu

STEPS = 4
CASES = [("fallback", negative_pressure_patch, MC), ("degenerate", cold_slow_flow, MINMOD)]


def oracle_steps(oracle, g, U0, lim, steps=STEPS):
    p = make_params(t_end=1e300, limiter=lim, max_steps=steps)
    return oracle.solve(g, U0, p, allow_max_steps=True, dt_cap=16)


def single_run(hip, U0, nx, ny, lim, precision, steps=STEPS):
    s = hip.MHDSolver(nx, ny)
    s.set_backend("hip", precision)
    if lim == MC:
        s.set_limiter_mc()
    else:
        s.set_limiter_minmod()
    s.state[...] = U0
    dev = s._dev()
    dev.set_dt_log(16)
    r = s.advance_steps(steps)
    return s.state.copy(), r, dev.dt_log(16)


@pytest.mark.parametrize("name,make,lim", CASES, ids=[c[0] for c in CASES])
def test_rare_branch_single_run_fp64_bitexact(hip, oracle, name, make, lim):
    g = make_grid(48, 40)
    U0 = make(oracle, g)
    Uo, info = oracle_steps(oracle, g, U0, lim)
    assert np.isfinite(Uo).all()
    U, r, dts = single_run(hip, U0, 48, 40, lim, "fp64")
    assert r.steps == STEPS and r.t == info["t"] and r.c_h == info["c_h"]
    np.testing.assert_array_equal(dts, info["dt"])
    assert np.array_equal(U, Uo), f"{name}: max |diff| {np.abs(U - Uo).max():.3e}"
    if name == "degenerate":
        assert 0.0 < r.c_h < 5e-15


@pytest.mark.parametrize("name,make,lim", CASES, ids=[c[0] for c in CASES])
def test_rare_branch_batch_fp64_bitexact(hip, oracle, name, make, lim):
    """The batched kernels on the same state, beside an ordinary instance (the branch is taken
    by the waves of the rare state only), each equal to the oracle bit for bit."""
    g = make_grid(48, 40)
    U0 = make(oracle, g)
    Uo, info = oracle_steps(oracle, g, U0, lim)
    g2 = make_grid(70, 33)
    V0 = oracle.init_ic("orszag_tang", g2)
    Vo, _ = oracle_steps(oracle, g2, V0, lim)
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter(lim), hbm_budget=1 << 28)
    i = eng.add_steps(U0, STEPS)
    j = eng.add_steps(V0, STEPS)
    eng.advance()
    r, fin = eng.result(i)
    assert fin and r.steps == STEPS and r.t == info["t"] and r.c_h == info["c_h"]
    assert np.array_equal(eng.download(i), Uo), name
    assert np.array_equal(eng.download(j), Vo)


def test_fallback_fp32_within_tolerance(hip, oracle):
    g = make_grid(48, 40)
    U0 = negative_pressure_patch(oracle, g)
    Uo, info = oracle_steps(oracle, g, U0, MC)
    outs = [single_run(hip, U0, 48, 40, MC, "fp32")[0]]
    eng = hip.BatchEngine("fp32", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    i = eng.add_steps(U0, STEPS)
    eng.advance()
    outs.append(eng.download(i))
    for U in outs:
        assert np.isfinite(U).all()
        for c in range(8):
            ref = Uo[c, 2:-2, 2:-2]
            den = np.abs(ref).sum()
            if den > 0:
                err = np.abs(U[c, 2:-2, 2:-2] - ref).sum() / den
                assert err <= 1e-4, (c, err)
