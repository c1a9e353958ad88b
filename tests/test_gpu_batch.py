"""Batched engine: every instance gives the oracle's fp64 bits, its result is the same alone
and in any mix (fp32 included), admission honours the HBM budget, and tiling does not matter."""

import numpy as np
import pytest

from checkers import MC, MINMOD, OUTFLOW, PERIODIC, make_grid, make_params

pytestmark = pytest.mark.gpu


def mixed_specs(hip, count=9, seed=3):
    from mhd_gpu_solver_amd import workloads
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        n = int(rng.choice([17, 40, 61, 64, 96, 130]))
        ot = k % 2 == 0
        s = workloads.Spec("orszag_tang_b0" if ot else "rotor_b0", n,
                           float(rng.uniform(1.4, 5 / 3)), float(rng.uniform(0.5, 2.0)),
                           (PERIODIC, PERIODIC) if ot else (OUTFLOW, PERIODIC),
                           t_end=float(rng.uniform(0.005, 0.02)))
        specs.append(s)
    return specs


def oracle_run(oracle, spec, U0, limiter, cfl=0.4, steps=None):
    g = make_grid(spec.n, spec.n)
    p = make_params(t_end=spec.t_end if steps is None else 1e300, cfl=cfl, gamma=spec.gamma,
                    limiter=limiter, max_steps=steps if steps is not None else 1_000_000)
    return oracle.solve(g, U0, p, bc=spec.bc, allow_max_steps=steps is not None)


def test_batch_fp64_matches_oracle_bitwise(hip, oracle):
    specs = mixed_specs(hip)
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 30)
    ids = []
    states = []
    for s in specs:
        U0 = s.state()
        states.append(U0)
        ids.append(eng.add(U0, s.t_end, gamma=s.gamma, bc=s.bc))
    assert eng.advance() == 0
    for s, i, U0 in zip(specs, ids, states):
        r, fin = eng.result(i)
        assert fin
        Uo, info = oracle_run(oracle, s, U0, MC)
        assert r.steps == info["steps"] and r.t == info["t"] and r.c_h == info["c_h"]
        assert np.array_equal(eng.download(i), Uo), (s.ic, s.n)


def test_batch_fp64_fixed_steps_and_single_run_agree(hip, oracle):
    s = mixed_specs(hip, 1, seed=11)[0]
    U0 = s.state()
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.Minmod, hbm_budget=1 << 28)
    i = eng.add_steps(U0, 13, gamma=s.gamma, bc=s.bc)
    eng.advance()
    r, fin = eng.result(i)
    assert fin and r.steps == 13
    Uo, info = oracle_run(oracle, s, U0, MINMOD, steps=13)
    assert np.array_equal(eng.download(i), Uo)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_same_alone_and_in_the_mix(hip, precision):
    """Continuous batching must not change a result: instance X alone, then X admitted into a
    running mix (others arriving before and after it) -> identical bits."""
    specs = mixed_specs(hip, 7, seed=5)
    x = specs[3]
    Ux = x.state()
    alone = hip.BatchEngine(precision, hbm_budget=1 << 30)
    ia = alone.add(Ux, x.t_end, gamma=x.gamma, bc=x.bc)
    alone.advance()
    ref = alone.download(ia)
    ref_r, _ = alone.result(ia)

    mix = hip.BatchEngine(precision, hbm_budget=1 << 30)
    early = [mix.add(s.state(), s.t_end, gamma=s.gamma, bc=s.bc) for s in specs[:3]]
    mix.advance(max_replays=2)                       # others are already running
    ix = mix.add(Ux, x.t_end, gamma=x.gamma, bc=x.bc)
    mix.advance(max_replays=1)
    late = [mix.add(s.state(), s.t_end, gamma=s.gamma, bc=s.bc) for s in specs[4:]]
    mix.remove(early[0])                             # a departure mid-run
    mix.advance()
    r, fin = mix.result(ix)
    assert fin and r == ref_r
    assert np.array_equal(mix.download(ix), ref)
    for i in early[1:] + late:
        assert mix.result(i)[1]


def test_admission_by_hbm_budget(hip):
    per = hip.instance_bytes("fp32", 64, 64)
    eng = hip.BatchEngine("fp32", hbm_budget=3 * per, max_instances=8)
    U = hip.init_ic("orszag_tang", 64, 64)
    ids = [eng.add_steps(U, 4) for _ in range(3)]
    with pytest.raises(hip.CapacityFull):
        eng.add_steps(U, 4)
    eng.remove(ids[1])
    ids.append(eng.add_steps(U, 4))
    st = eng.stats()
    assert st["instances"] == 3 and st["bytes_used"] <= st["bytes_budget"]
    eng.advance()
    outs = [eng.download(i) for i in (ids[0], ids[2], ids[3])]
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


@pytest.mark.parametrize("rows", [0, -1])
def test_batch_prepare_then_advance_matches_oracle(hip, oracle, rows):
    """prepare() plans the next replay (tables, timed row choice, graph) without stepping: on
    an empty engine it does nothing, called twice it changes nothing, and the solves that
    follow equal the oracle bit for bit (state, steps, t)."""
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 30)
    eng.set_seg_rows(rows)
    eng.prepare()
    specs = mixed_specs(hip, 4, seed=29)
    runs = [(eng.add(s.state(), s.t_end, gamma=s.gamma, bc=s.bc), s) for s in specs]
    eng.prepare()
    eng.prepare()
    for i, s in runs:
        assert eng.result(i)[0].steps == 0
    assert eng.advance() == 0
    for i, s in runs:
        Uo, info = oracle_run(oracle, s, s.state(), MC)
        r, fin = eng.result(i)
        assert fin and r.steps == info["steps"] and r.t == info["t"]
        assert np.array_equal(eng.download(i), Uo)


@pytest.mark.parametrize("rows", [1, 5, 200, -1])
def test_batch_tiling_independent(hip, oracle, rows):
    s = mixed_specs(hip, 1, seed=17)[0]
    U0 = s.state()
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    eng.set_seg_rows(rows)
    i = eng.add(U0, s.t_end, gamma=s.gamma, bc=s.bc)
    eng.advance()
    Uo, info = oracle_run(oracle, s, U0, MC)
    assert eng.result(i)[0].steps == info["steps"]
    assert np.array_equal(eng.download(i), Uo)


@pytest.mark.parametrize("limiter", [MINMOD, MC])
def test_batch_nonideal_rectangular_fp64_matches_oracle(hip, oracle, limiter):
    """All three non-ideal terms (the batch's NI kernels), rectangular grids, offset domains and
    mixed boundaries: every instance equals the oracle bit for bit (state, steps, t, c_h)."""
    ni = hip.NonIdealConfig(ohmic=True, hall=True, ambipolar=True, eta_ohm=1e-3,
                            eta_hall=1e-2, eta_ambipolar=1e-3)
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter(limiter), nonideal=ni,
                          hbm_budget=1 << 28)
    cases = [("rotor", 70, 45, (OUTFLOW, OUTFLOW), (-0.5, 0.5, -0.25, 0.5)),
             ("orszag_tang", 33, 61, (PERIODIC, PERIODIC), (0.0, 1.0, 0.0, 2.0)),
             ("rotor", 128, 20, (PERIODIC, OUTFLOW), (0.0, 2.0, 0.0, 1.0))]
    runs = []
    for name, nx, ny, bc, (x0, x1, y0, y1) in cases:
        g = make_grid(nx, ny, x0, x1, y0, y1)
        U0 = oracle.init_ic(name, g)
        i = eng.add(U0, 0.004, gamma=5.0 / 3.0, bc=bc, x_min=x0, x_max=x1, y_min=y0, y_max=y1)
        runs.append((i, g, U0, bc))
    eng.advance()
    for i, g, U0, bc in runs:
        r, fin = eng.result(i)
        Uo, info = oracle.solve(g, U0, make_params(t_end=0.004, limiter=limiter, eta_ohm=1e-3,
                                                   eta_hall=1e-2, eta_ambipolar=1e-3), bc=bc)
        assert fin and r.steps == info["steps"] and r.t == info["t"] and r.c_h == info["c_h"]
        assert np.array_equal(eng.download(i), Uo), (g.nx, g.ny, bc)


@pytest.mark.parametrize("rows", [0, 1, 2, 200, -1])
def test_batch_half_wave_tiles_edges_match_oracle(hip, oracle, rows):
    """Narrow tail strips (at most 28 columns) march on half-wave tiles in the fp64 batch: the
    lower and upper halves of a wave take the same chunk of the strip's lower and upper row
    halves. Edges: tails of 1 and 28 columns, a grid of one tail strip only, 1 to 3 rows (the
    upper half then updates one row or none), odd and even row counts, every boundary mix and
    rows per wave from 1 to more than the grid (0: the load model's choice; -1: steps taken in
    turn with the timed shortlist's partitions)."""
    eng = hip.BatchEngine("fp64", limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
    if rows:
        eng.set_seg_rows(rows)
    cases = [("orszag_tang", 61, 1, (PERIODIC, PERIODIC)),
             ("rotor", 88, 2, (OUTFLOW, OUTFLOW)),
             ("orszag_tang", 28, 3, (PERIODIC, OUTFLOW)),
             ("rotor", 121, 9, (OUTFLOW, PERIODIC)),
             ("orszag_tang", 5, 40, (PERIODIC, PERIODIC))]
    runs = []
    for name, nx, ny, bc in cases:
        g = make_grid(nx, ny, 0.0, 1.0, 0.0, 1.0)
        U0 = oracle.init_ic(name, g)
        runs.append((eng.add_steps(U0, 7, bc=bc), g, U0, bc))
    eng.advance()
    for i, g, U0, bc in runs:
        r, fin = eng.result(i)
        Uo, info = oracle.solve(g, U0, make_params(t_end=1e300, limiter=MC, max_steps=7), bc=bc,
                                allow_max_steps=True)
        assert fin and r.steps == 7 and r.t == info["t"] and r.c_h == info["c_h"]
        assert np.array_equal(eng.download(i), Uo), (g.nx, g.ny, bc, rows)


@pytest.mark.parametrize("precision", ["fp32", "fp64_fast"])
def test_batch_fp32_narrow_grids_tiling_independent_and_within_l1(hip, oracle, precision):
    """The fp32 engine (and the contracted fp64 one, whose odd bands march the other way) on
    narrow and short grids (tails of 1 to 28 columns, 1 to 3 rows): the same bits for every
    rows-per-wave choice, and within the fp32 bar (1e-4 relative L1 per field) of the
    oracle's fp64 result."""
    cases = [("orszag_tang", 61, 1, (PERIODIC, PERIODIC)),
             ("rotor", 88, 2, (OUTFLOW, OUTFLOW)),
             ("orszag_tang", 28, 3, (PERIODIC, OUTFLOW)),
             ("rotor", 121, 9, (OUTFLOW, PERIODIC)),
             ("orszag_tang", 64, 40, (PERIODIC, PERIODIC))]
    states = [oracle.init_ic(name, make_grid(nx, ny, 0.0, 1.0, 0.0, 1.0))
              for name, nx, ny, _ in cases]
    first = None
    for rows in (0, 1, 2, 200, -1):
        eng = hip.BatchEngine(precision, limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
        if rows:
            eng.set_seg_rows(rows)
        ids = [eng.add_steps(U0, 7, bc=c[3]) for U0, c in zip(states, cases)]
        eng.advance()
        outs = [eng.download(i) for i in ids]
        if first is None:
            first = outs
        for a, b, c in zip(outs, first, cases):
            assert np.array_equal(a, b), (c, rows)
    for out, U0, (name, nx, ny, bc) in zip(first, states, cases):
        g = make_grid(nx, ny, 0.0, 1.0, 0.0, 1.0)
        Uo, _ = oracle.solve(g, U0, make_params(t_end=1e300, limiter=MC, max_steps=7), bc=bc,
                             allow_max_steps=True)
        for c in range(8):
            ref = Uo[c, 2:-2, 2:-2]
            den = np.abs(ref).sum()
            if den > 0:
                assert np.abs(out[c, 2:-2, 2:-2] - ref).sum() / den <= 1e-4, (name, nx, ny, c)


def test_batch_engines_in_threads_equal_alone(hip):
    """Two batch engines (fp32 and fp64, each with its own stream and captured graph) advanced
    from two Python threads at once give the same bits and step counts as each advanced alone."""
    import threading

    specs = mixed_specs(hip, 6, seed=23)
    states = [s.state() for s in specs]

    def run(precision):
        eng = hip.BatchEngine(precision, limiter=hip.SlopeLimiter.MC, hbm_budget=1 << 28)
        ids = [eng.add(U, s.t_end, gamma=s.gamma, bc=s.bc) for s, U in zip(specs, states)]
        eng.advance()
        return [(eng.result(i)[0], eng.download(i)) for i in ids]

    alone = {p: run(p) for p in ("fp32", "fp64")}
    out, errors = {}, []

    def work(p):
        try:
            out[p] = run(p)
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(p,)) for p in ("fp32", "fp64")]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for p in ("fp32", "fp64"):
        assert p in out
        for (ra, Ua), (rb, Ub) in zip(alone[p], out[p]):
            assert ra == rb
            assert np.array_equal(Ua, Ub)


def test_batch_fp32_nonideal_tails_tiling_independent_and_within_l1(hip, oracle):
    """The fp32 non-ideal batch kernel on grids with narrow tail strips (half-wave tiles) and
    offset rectangular domains: the same bits for 1 and 200 rows per wave and the load model's
    choice, and within the fp32 bar (1e-4 relative L1 per field) of the oracle's fp64 result."""
    ni = hip.NonIdealConfig(ohmic=True, hall=True, ambipolar=True, eta_ohm=1e-3,
                            eta_hall=1e-2, eta_ambipolar=1e-3)
    cases = [("rotor", 70, 45, (OUTFLOW, OUTFLOW), (-0.5, 0.5, -0.25, 0.5)),
             ("orszag_tang", 88, 33, (PERIODIC, PERIODIC), (0.0, 1.0, 0.0, 2.0)),
             ("rotor", 128, 20, (PERIODIC, OUTFLOW), (0.0, 2.0, 0.0, 1.0))]
    grids = [make_grid(nx, ny, x0, x1, y0, y1) for _, nx, ny, _, (x0, x1, y0, y1) in cases]
    states = [oracle.init_ic(c[0], g) for c, g in zip(cases, grids)]
    first = None
    for rows in (0, 1, 200):
        eng = hip.BatchEngine("fp32", limiter=hip.SlopeLimiter.MC, nonideal=ni,
                              hbm_budget=1 << 28)
        if rows:
            eng.set_seg_rows(rows)
        ids = [eng.add_steps(U0, 6, gamma=5.0 / 3.0, bc=c[3], x_min=c[4][0], x_max=c[4][1],
                             y_min=c[4][2], y_max=c[4][3])
               for U0, c in zip(states, cases)]
        eng.advance()
        outs = [eng.download(i) for i in ids]
        if first is None:
            first = outs
        for a, b, c in zip(outs, first, cases):
            assert np.array_equal(a, b), (c[:3], rows)
    for out, U0, g, c in zip(first, states, grids, cases):
        Uo, _ = oracle.solve(g, U0, make_params(t_end=1e300, limiter=MC, eta_ohm=1e-3,
                                                eta_hall=1e-2, eta_ambipolar=1e-3,
                                                max_steps=6),
                             bc=c[3], allow_max_steps=True)
        for k in range(8):
            ref = Uo[k, 2:-2, 2:-2]
            den = np.abs(ref).sum()
            if den > 0:
                assert np.abs(out[k, 2:-2, 2:-2] - ref).sum() / den <= 1e-4, (c[:3], k)
