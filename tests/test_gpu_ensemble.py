"""The 512-instance mixed-size ensemble (workloads.ensemble(512, seed 2024): Orszag-Tang /
rotor, N in {64, 96, 128, 192, 256}, gamma ~ U[1.4, 5/3], field scale ~ U[0.5, 2], minmod,
CFL 0.4) on the batched engine at its full size, with the engine's default work partition:

* fp64: every instance steps 8 times; a sample holding every (size, initial condition) pair
  several times equals the oracle bit for bit (state, t, steps, c_h), following the per-instance
  solve contract (include/mhd/time_integrator.hpp:62-106);
* FP32-bulk: all 512 instances within 1e-4 relative L1 per conserved field of the fp64 batch
  (computed on the device), with equal step counts."""
import collections

import numpy as np
import pytest

from checkers import MINMOD, make_grid, make_params

pytestmark = pytest.mark.gpu

STEPS = 8
PER_PAIR = 5   # oracle samples per (size, initial condition)


@pytest.fixture(scope="module")
def ensemble(hip):
    from mhd_gpu_solver_amd import workloads
    specs = workloads.ensemble(512, 2024)
    states = [s.state() for s in specs]
    engines = {}
    for prec in ("fp64", "fp32"):
        budget = sum(hip.instance_bytes(prec, s.n, s.n) for s in specs) + (64 << 20)
        eng = hip.BatchEngine(prec, limiter=hip.SlopeLimiter.Minmod, hbm_budget=budget,
                              max_instances=len(specs))
        ids = [eng.add_steps(U, STEPS, gamma=s.gamma, bc=s.bc) for s, U in zip(specs, states)]
        assert eng.advance() == 0
        engines[prec] = (eng, ids)
    return specs, states, engines


def test_ensemble_covers_the_configuration(ensemble):
    specs, _, _ = ensemble
    pairs = collections.Counter((s.n, s.ic) for s in specs)
    assert len(specs) == 512
    assert {n for n, _ in pairs} == {64, 96, 128, 192, 256}
    assert {ic for _, ic in pairs} == {"orszag_tang_b0", "rotor_b0"}
    assert min(pairs.values()) >= PER_PAIR


def test_ensemble_fp64_sample_bitexact_vs_oracle(ensemble, oracle):
    specs, states, engines = ensemble
    eng, ids = engines["fp64"]
    taken = collections.Counter()
    checked = 0
    for k, s in enumerate(specs):
        key = (s.n, s.ic)
        if taken[key] >= PER_PAIR:
            continue
        taken[key] += 1
        r, fin = eng.result(ids[k])
        assert fin and r.steps == STEPS
        p = make_params(t_end=1e300, cfl=0.4, gamma=s.gamma, limiter=MINMOD, max_steps=STEPS)
        Uo, info = oracle.solve(make_grid(s.n, s.n), states[k], p, bc=s.bc,
                                allow_max_steps=True)
        assert r.t == info["t"] and r.c_h == info["c_h"], (k, s.n, s.ic)
        assert np.array_equal(eng.download(ids[k]), Uo), (k, s.n, s.ic)
        checked += 1
    assert checked == 10 * PER_PAIR


def test_ensemble_fp32_all_within_l1_of_fp64(ensemble):
    specs, _, engines = ensemble
    e64, ids64 = engines["fp64"]
    e32, ids32 = engines["fp32"]
    worst = np.zeros(9)
    for k, s in enumerate(specs):
        r32, f32 = e32.result(ids32[k])
        r64, f64 = e64.result(ids64[k])
        assert f32 and f64 and r32.steps == r64.steps == STEPS
        l1 = e32.rel_l1(ids32[k], e64, ids64[k])
        worst = np.maximum(worst, l1)
    print("ensemble fp32 vs fp64 worst rel L1 per field:", " ".join(f"{v:.2e}" for v in worst))
    assert worst[:8].max() <= 1e-4, worst
