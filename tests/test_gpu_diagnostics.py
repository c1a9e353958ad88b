"""On-device compensated totals (mass/energy drift) and drift bounds of the reference's GPU
gate (docs/gpu.md:129-137: drift <= 1e-11 fp64, <= 2e-6 fp32)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fsum_totals(U):
    return np.array([math.fsum(U[c, 2:-2, 2:-2].ravel().tolist()) for c in range(9)])


def assert_dd_sum(got, U):
    """Double-double summation of n terms: |error| <= ~n u^2 sum|x| before the final rounding
    (u = 2^-53), then one rounding; checked against the exactly rounded math.fsum."""
    want = fsum_totals(U)
    absum = np.abs(U[:, 2:-2, 2:-2]).reshape(9, -1).sum(axis=1)
    n = U[0, 2:-2, 2:-2].size
    bound = 2.3e-16 * np.abs(want) + 4 * n * 2.0 ** -106 * absum + 1e-300
    assert np.all(np.abs(got - want) <= bound), (got - want, bound)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_totals_match_compensated_bound(hip, precision):
    s = hip.MHDSolver(173, 91)
    s.set_backend("hip", precision)
    s.initialize("rotor")
    s.run(0.01)
    dev = s._dev()
    dev.upload(s.state)
    got = dev.totals()
    assert_dd_sum(got, s.state)                # s.state holds the device values exactly
    assert np.array_equal(dev.totals(), got)   # deterministic


@pytest.mark.parametrize("precision,bound", [("fp64", 1e-11), ("fp32", 2e-6)])
def test_mass_energy_drift_periodic(hip, precision, bound):
    s = hip.MHDSolver(128, 128)
    s.set_backend("hip", precision)
    s.set_limiter_mc()
    s.initialize("orszag_tang")
    t0 = s.conserved_totals()
    s.run(0.1)
    t1 = s.conserved_totals()
    for c, name in ((0, "mass"), (4, "energy")):
        drift = abs(t1[c] - t0[c]) / abs(t0[c])
        assert drift <= bound, (name, drift)


def test_batch_totals_match_single_run(hip):
    from mhd_gpu_solver_amd import workloads
    spec = workloads.ensemble(3, seed=1)[0]
    U0 = spec.state()
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.Minmod, hbm_budget=1 << 28)
    i = eng.add_steps(U0, 16, gamma=spec.gamma, bc=spec.bc)
    eng.advance()
    assert_dd_sum(eng.totals(i), eng.download(i))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_device_max_div_b_equals_oracle(hip, oracle, precision):
    """Device max |div B| after a solve equals the oracle's on the downloaded state, bit for
    bit (fp64 arithmetic from the stored values in both precisions); on the OT 32^2 golden
    case it also reproduces the golden's recorded value."""
    from checkers import make_grid
    s = hip.MHDSolver(32, 32)
    s.set_backend("hip", precision)
    s.set_limiter_mc()
    s.initialize("orszag_tang")
    s.run(0.05)
    dev = s._dev()
    got = dev.max_div_b()
    assert got == oracle.max_div_b(make_grid(32, 32), s.state)
    assert got == s.max_div_b()
    if precision == "fp64":
        assert abs(got - 1.713604) <= 5e-7   # tests/goldens/orszag_tang_32.golden: max_div_b


def test_batch_max_div_b_equals_oracle(hip, oracle):
    from checkers import make_grid
    from mhd_gpu_solver_amd import workloads
    eng = hip.BatchEngine("fp32", hbm_budget=1 << 28)
    specs = [workloads.Spec("rotor", 48, 5.0 / 3.0, 1.0, (2, 2), 0.01),
             workloads.Spec("kh_tanh", 40, 5.0 / 3.0, 1.0, (0, 0), 0.01, seed=3)]
    ids = [eng.add(sp.state(), sp.t_end, gamma=sp.gamma, bc=sp.bc) for sp in specs]
    eng.advance()
    vals = []
    for i, sp in zip(ids, specs):
        vals.append(eng.max_div_b(i))
        assert vals[-1] == oracle.max_div_b(make_grid(sp.n, sp.n), eng.download(i))
    assert max(vals) > 0.0


def rel_l1_exact(A, B):
    """sum |a - b| / sum |b| over the interior per component (the absolute sum where sum |b| is
    0), with the differences taken in fp64 from the stored values and exactly rounded sums."""
    out = []
    for c in range(9):
        a = A[c, 2:-2, 2:-2].ravel()
        b = B[c, 2:-2, 2:-2].ravel()
        d = math.fsum(np.abs(a - b).tolist())
        r = math.fsum(np.abs(b).tolist())
        out.append(d / r if r > 0 else d)
    return np.array(out)


def test_rel_l1_on_device_fp32_against_fp64(hip):
    """The device comparison of an FP32 state with an fp64 state of the same solve equals the
    exactly rounded host computation to double-double accuracy, is zero for a state against
    itself, gives the same bits on every call and refuses states of different grids."""
    states = {}
    for prec in ("fp32", "fp64"):
        s = hip.MHDSolver(173, 91)
        s.set_backend("hip", prec)
        s.initialize("rotor")
        s.run(0.01)
        dev = s._dev()
        dev.upload(s.state)          # the state the host holds, exactly representable
        states[prec] = (s, dev)
    (s32, d32), (s64, d64) = states["fp32"], states["fp64"]
    got = d32.rel_l1(d64)
    want = rel_l1_exact(s32.state, s64.state)
    assert np.allclose(got, want, rtol=1e-13, atol=0.0), (got, want)
    assert max(got[:8]) <= 1e-4
    assert np.array_equal(d32.rel_l1(d64), got)
    assert np.array_equal(d64.rel_l1(d64), np.zeros(9))
    other = hip.MHDSolver(172, 91)
    other.set_backend("hip", "fp64")
    other.initialize("rotor")
    with pytest.raises(ValueError):
        d64.rel_l1(other._dev())


def test_batch_rel_l1_across_engines(hip):
    """Instance against instance across an FP32 and an fp64 engine, on the device, equals the
    host computation from the downloaded states."""
    from mhd_gpu_solver_amd import workloads
    spec = workloads.Spec("orszag_tang_b0", 130, 1.5, 1.2, (0, 0), t_end=0.02)
    U0 = spec.state()
    engines, ids = {}, {}
    for prec in ("fp32", "fp64"):
        eng = hip.BatchEngine(prec, limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
        ids[prec] = eng.add(U0, spec.t_end, gamma=spec.gamma, bc=spec.bc)
        eng.advance()
        engines[prec] = eng
    got = engines["fp32"].rel_l1(ids["fp32"], engines["fp64"], ids["fp64"])
    want = rel_l1_exact(engines["fp32"].download(ids["fp32"]),
                        engines["fp64"].download(ids["fp64"]))
    assert np.allclose(got, want, rtol=1e-13, atol=0.0), (got, want)
    assert np.array_equal(engines["fp64"].rel_l1(ids["fp64"], engines["fp64"], ids["fp64"]),
                          np.zeros(9))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stream_bandwidth_is_a_plausible_rate(hip, precision):
    """The streaming measurement beside the roofline: positive, and below the published HBM
    peak of 8 TB/s (plus timer slack) for both stage mixes."""
    for stage in (1, 2):
        gbs = hip.stream_bandwidth(precision, 4_000_000, stage, reps=5)
        assert 100.0 < gbs < 8000.0 * 1.05
