"""Continuous batching: requests arriving while others run are admitted as memory frees up,
and each result is identical to solving the same request alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_stream(seed=9, count=10):
    from mhd_gpu_solver_amd import workloads
    reqs = workloads.request_stream(count, rate=400.0, seed=seed, n_min=24, n_max=72,
                                    t_min=0.01, t_max=0.03)
    return reqs


def alone(hip, req, precision="fp32"):
    eng = hip.BatchEngine(precision, limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    i = eng.add(req.state(), req.t_end, gamma=req.gamma, bc=req.bc)
    eng.advance()
    return eng.download(i), eng.result(i)[0]


@pytest.mark.parametrize("budget_instances", [64, 2])
def test_served_results_equal_alone(hip, budget_instances):
    from mhd_gpu_solver_amd.serve import Server
    reqs = small_stream()
    per = max(hip.instance_bytes("fp32", r.n, r.n) for r in reqs)
    srv = Server("fp32", hbm_budget=budget_instances * per, max_instances=64)
    rep = srv.run(reqs, realtime=True, keep_states=True)
    assert len(rep.completed) == len(reqs)
    assert rep.max_concurrent <= budget_instances
    if budget_instances == 2:
        assert rep.max_concurrent == 2
    for c in rep.completed:
        ref, r = alone(hip, reqs[c.index])
        assert c.steps == r.steps and c.t == r.t
        assert np.array_equal(c.state, ref), (c.index, reqs[c.index].ic)
        assert c.finished >= c.admitted >= c.arrival
    s = rep.summary()
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_served_fp64_matches_oracle(hip, oracle):
    from checkers import MC, make_grid, make_params
    from mhd_gpu_solver_amd.serve import Server
    reqs = small_stream(seed=4, count=4)
    srv = Server("fp64", hbm_budget=1 << 30, max_instances=8)
    rep = srv.run(reqs, realtime=False, keep_states=True)
    for c in rep.completed:
        req = reqs[c.index]
        Uo, info = oracle.solve(make_grid(req.n, req.n), req.state(),
                                make_params(t_end=req.t_end, gamma=req.gamma, limiter=MC),
                                bc=req.bc)
        assert c.steps == info["steps"]
        assert np.array_equal(c.state, Uo)
