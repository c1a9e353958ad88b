"""The batched engine with per-instance schemes and geometry: each instance has its own slope
limiter, non-ideal terms (Ohmic / Hall / ambipolar coefficients), uint8 FluidMask and
reflective-wall setting (include/mhd/mask.hpp:17-165, include/mhd/rhs.hpp:116-240), and
instances of different schemes advance together in one engine. fp64: every instance equals the
oracle bit for bit (state, t, steps, c_h), following the reference's solve contract per
instance (include/mhd/time_integrator.hpp:62-106); results are the same alone and in any mix;
the non-ideal + geometry workload (rotor 1024^2, eight solid circles) matches the single-run
solver; a served masked request equals the oracle."""
import numpy as np
import pytest

from checkers import MC, MINMOD, OUTFLOW, PERIODIC, make_grid, make_params

pytestmark = pytest.mark.gpu


def edge_mask(nx, ny):
    """Solids on the domain edges (stale-ghost semantics), a one-cell solid, a block."""
    m = np.ones((ny, nx), dtype=np.uint8)
    m[0:5, 0:3] = 0
    m[ny - 6:ny, nx - 4:nx] = 0
    m[ny // 3, nx // 2] = 0
    m[ny // 2:ny // 2 + 3, 5:8] = 0
    return m


def block_mask(nx, ny):
    m = np.ones((ny, nx), dtype=np.uint8)
    m[ny // 3:ny // 3 + 8, nx // 3:nx // 3 + 6] = 0
    return m


def cases(hip):
    """(ic, nx, ny, bc, limiter, nonideal dict, mask, walls)"""
    return [
        ("orszag_tang", 48, 40, (PERIODIC, PERIODIC), MC, {}, None, True),
        ("rotor", 44, 36, (OUTFLOW, OUTFLOW), MINMOD,
         dict(eta_ohm=1e-3, eta_hall=1e-2, eta_ambipolar=1e-3), edge_mask(44, 36), True),
        ("orszag_tang", 40, 40, (PERIODIC, PERIODIC), MC, {}, block_mask(40, 40), True),
        ("rotor", 50, 30, (OUTFLOW, PERIODIC), MC, dict(eta_ohm=2e-3), block_mask(50, 30),
         False),
        ("orszag_tang", 64, 64, (PERIODIC, PERIODIC), MINMOD, {}, None, True),
        ("rotor", 61, 33, (PERIODIC, OUTFLOW), MINMOD, dict(eta_hall=5e-3), None, True),
    ]


def nonideal_config(hip, ni):
    return hip.NonIdealConfig(ohmic="eta_ohm" in ni, hall="eta_hall" in ni,
                              ambipolar="eta_ambipolar" in ni, eta_ohm=ni.get("eta_ohm", 0.0),
                              eta_hall=ni.get("eta_hall", 0.0),
                              eta_ambipolar=ni.get("eta_ambipolar", 0.0))


def admit(hip, eng, oracle, case, t_end):
    ic, nx, ny, bc, lim, ni, mask, walls = case
    g = make_grid(nx, ny)
    U0 = oracle.init_ic(ic, g, bc=bc)
    if lim == MC and not ni and mask is None:
        i = eng.add(U0, t_end, bc=bc)          # the engine's own scheme (MC, ideal)
    else:
        i = eng.add(U0, t_end, bc=bc, limiter=hip.SlopeLimiter(lim),
                    nonideal=nonideal_config(hip, ni), mask=mask, walls=walls)
    return i, g, U0


def oracle_case(oracle, case, g, U0, t_end):
    ic, nx, ny, bc, lim, ni, mask, walls = case
    p = make_params(t_end=t_end, limiter=lim, eta_ohm=ni.get("eta_ohm"),
                    eta_hall=ni.get("eta_hall"), eta_ambipolar=ni.get("eta_ambipolar"))
    return oracle.solve(g, U0, p, bc=bc, mask=mask, walls=walls)


@pytest.mark.parametrize("rows", [0, 3, -1])
def test_mixed_schemes_and_masks_fp64_bitexact(hip, oracle, rows):
    """rows -1: the timed choice's dry steps (every variant, the masked fill) change nothing."""
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    if rows:
        eng.set_seg_rows(rows)
    t_end = 0.015
    runs = [(c,) + admit(hip, eng, oracle, c, t_end) for c in cases(hip)]
    assert eng.advance() == 0
    for case, i, g, U0 in runs:
        r, fin = eng.result(i)
        Uo, info = oracle_case(oracle, case, g, U0, t_end)
        assert fin and r.steps == info["steps"] and r.t == info["t"] and r.c_h == info["c_h"], \
            case[:3]
        U = eng.download(i)
        assert np.array_equal(U, Uo), (case[:3], np.abs(U - Uo).max())
    assert eng.stats()["instances"] == len(runs)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_masked_instance_same_alone_and_in_mix(hip, oracle, precision):
    cs = cases(hip)
    x = cs[1]
    t_end = 0.01
    alone = hip.BatchEngine(precision, limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    ia, _, _ = admit(hip, alone, oracle, x, t_end)
    alone.advance()
    ref = alone.download(ia)
    ref_r, _ = alone.result(ia)
    mix = hip.BatchEngine(precision, limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    early = [admit(hip, mix, oracle, c, t_end)[0] for c in cs[2:4]]
    mix.advance(max_replays=1)
    ix, _, _ = admit(hip, mix, oracle, x, t_end)
    late = [admit(hip, mix, oracle, c, t_end)[0] for c in (cs[0], cs[4], cs[5])]
    mix.remove(early[0])
    mix.advance()
    r, fin = mix.result(ix)
    assert fin and r == ref_r
    assert np.array_equal(mix.download(ix), ref)
    for i in early[1:] + late:
        assert mix.result(i)[1]


def test_fp32_mixed_schemes_within_tolerance(hip, oracle):
    eng = hip.BatchEngine("fp32", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    t_end = 0.015
    runs = [(c,) + admit(hip, eng, oracle, c, t_end) for c in cases(hip)]
    eng.advance()
    for case, i, g, U0 in runs:
        Uo, info = oracle_case(oracle, case, g, U0, t_end)
        r, _ = eng.result(i)
        assert abs(r.steps - info["steps"]) <= 1
        U = eng.download(i)
        for c in range(8):
            ref = Uo[c, 2:-2, 2:-2]
            den = np.abs(ref).sum()
            if den > 0:
                err = np.abs(U[c, 2:-2, 2:-2] - ref).sum() / den
                assert err <= 1e-4, (case[:3], c, err)


def test_rotor_circles_1024_batched_equals_single_run(hip):
    """The non-ideal + geometry workload (rotor 1024^2, Ohmic 1e-3 + Hall 1e-2 + ambipolar
    1e-3, outflow, eight seeded solid circles as reflective walls) batched beside an ordinary
    instance, 12 steps, equals the single-run solver bit for bit (which equals the oracle:
    tests/test_gpu_configs.py)."""
    from mhd_gpu_solver_amd import workloads
    spec = workloads.rotor_circles(1024)
    U0 = spec.state()
    s = hip.MHDSolver(1024, 1024)
    s.set_backend("hip", "fp64")
    s.set_outflow_bc()
    s.set_limiter_minmod()
    s.enable_ohmic(1e-3)
    s.enable_hall(1e-2)
    s.enable_ambipolar(1e-3)
    s.state[...] = U0
    s.set_mask(spec.mask)
    U0 = s.state.copy()   # with the wall fill of set_mask, as the single run starts
    r1 = s.advance_steps(12)
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 31)
    i = eng.add_steps(U0, 12, bc=spec.bc, limiter=spec.limiter, nonideal=spec.nonideal,
                      mask=spec.mask, walls=True)
    j = eng.add_steps(hip.init_ic("orszag_tang", 200, 130), 12)
    eng.advance()
    r, fin = eng.result(i)
    assert fin and r.steps == 12 and r.t == r1.t
    assert np.array_equal(eng.download(i), s.state)
    assert eng.result(j)[1]


def test_served_masked_request_matches_oracle(hip, oracle):
    from mhd_gpu_solver_amd import workloads
    from mhd_gpu_solver_amd.serve import Server
    reqs = workloads.request_stream(5, rate=400.0, seed=3, n_min=24, n_max=64, t_min=0.01,
                                    t_max=0.02)
    masked = workloads.rotor_circles(96, seed=5, t_end=0.01)
    masked.arrival = reqs[2].arrival
    reqs.insert(3, masked)
    srv = Server("fp64", hbm_budget=1 << 28, max_instances=16)
    rep = srv.run(reqs, realtime=False, keep_states=True)
    assert len(rep.completed) == len(reqs)
    done = {c.index: c for c in rep.completed}
    c = done[3]
    assert c.error is None
    U0 = masked.state()
    p = make_params(t_end=0.01, limiter=MINMOD, eta_ohm=1e-3, eta_hall=1e-2,
                    eta_ambipolar=1e-3)
    # both start from the request's state as built; each stage's ghost and wall fill follows
    Uo, info = oracle.solve(make_grid(96, 96), U0, p, bc=masked.bc, mask=masked.mask,
                            walls=True)
    assert c.steps == info["steps"] and c.t == info["t"]
    assert np.array_equal(c.state, Uo), np.abs(c.state - Uo).max()
