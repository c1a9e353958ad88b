"""Pins of the CPU oracle (oracle/mhd_oracle.c) before it is trusted as the parity checker.

* Reference fixtures: the regression goldens (tests/goldens/*.golden of the reference, copied
  as data into tests/golden/) and the verification scalars (tests/verification/reference/*.ref),
  with the reference's own tolerances (tests/test_regression_golden.cpp:144-171,
  tests/test_verification_ci.cpp:58-131), and the published step count of Orszag-Tang 256^2 to
  t = 0.5 (docs/benchmarks.md:83).
* The reference compiled from its sources (oracle/_ref): bit-exact agreement on whole solves,
  dt sequences, single RHS evaluations, CFL, face fluxes, ghost and wall fills.
* Mathematics: limiter closed forms, HLL consistency F(W, W) = F(W), uniform-state invariance,
  conservation of the periodic totals, GLM flux terms, the non-ideal E field, and integer-tag
  brute force of the ghost indexing.
A planted mistake in any of these (sign, index, dropped term, transposed operand) fails a check.
"""
import math
import pathlib

import numpy as np
import pytest

from checkers import (FIELDS, MC, MINMOD, OUTFLOW, PERIODIC, empty_state, exact_reference,
                      interior, make_grid, make_params, read_golden, read_ref_scalars)

GOLD = pathlib.Path(__file__).parent / "golden"


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---- reference fixtures ------------------------------------------------------------------

def test_orszag_tang_32_golden(oracle):
    g = make_grid(32, 32)
    U0 = oracle.init_ic("orszag_tang", g)
    U, info = oracle.solve(g, U0, make_params(t_end=0.05, limiter=MC))
    meta, fields = read_golden(GOLD / "orszag_tang_32.golden")
    assert info["steps"] == int(meta["steps"])
    for name, ref in fields.items():
        got = interior(U)[FIELDS.index(name)].ravel()
        assert rel_l2(got, ref) <= 1e-8, name
        assert np.abs(got - ref).max() <= 1e-7, name


def test_max_div_b_golden_and_reference(oracle):
    """max |div B| after the OT 32^2 golden solve: the golden's recorded value (6 decimals),
    and bit-equal to the reference's max_abs_div_b when the reference build is present."""
    g = make_grid(32, 32)
    U, _ = oracle.solve(g, oracle.init_ic("orszag_tang", g), make_params(t_end=0.05, limiter=MC))
    meta, _ = read_golden(GOLD / "orszag_tang_32.golden")
    got = oracle.max_div_b(g, U)
    assert abs(got - float(meta["max_div_b"])) <= 5e-7
    U0 = oracle.init_ic("orszag_tang", g)
    # Bx depends on y only and By on x only: every central difference is exactly zero
    assert oracle.max_div_b(g, U0) == 0.0
    ref = exact_reference()
    if ref is not None:
        assert ref.max_div_b(g, U) == got


def test_ohmic_sine_bz_32_golden(oracle):
    g = make_grid(32, 32)
    U = empty_state(g)
    x = (np.arange(32) + 0.5) / 32
    bz = np.sin(2.0 * math.pi * x)
    U[0, 2:-2, 2:-2] = 1.0
    U[7, 2:-2, 2:-2] = bz[None, :]
    U[4, 2:-2, 2:-2] = 1.0 / (5.0 / 3.0 - 1.0) + 0.5 * bz[None, :] ** 2
    oracle.apply_bc(g, U, (PERIODIC, PERIODIC))
    U, info = oracle.solve(g, U, make_params(t_end=0.1, cfl=0.3, limiter=MINMOD, eta_ohm=0.1))
    meta, fields = read_golden(GOLD / "ohmic_sine_bz_32.golden")
    assert info["steps"] == int(meta["steps"])
    for name, ref in fields.items():
        got = interior(U)[FIELDS.index(name)].ravel()
        assert rel_l2(got, ref) <= 5e-8, name


def diagnostics(U, g, gamma):
    """include/verification/diagnostics.hpp:26-59"""
    Ui = interior(U)
    rho = Ui[0]
    ekin = 0.5 * (Ui[1] ** 2 + Ui[2] ** 2 + Ui[3] ** 2) / rho
    emag = 0.5 * (Ui[5] ** 2 + Ui[6] ** 2 + Ui[7] ** 2)
    p = (gamma - 1.0) * (Ui[4] - ekin - emag)
    dx = (g.x_max - g.x_min) / g.nx
    dy = (g.y_max - g.y_min) / g.ny
    return dict(min_rho=rho.min(), max_rho=rho.max(), min_pressure=p.min(),
                max_pressure=p.max(), mass=(rho * dx * dy).sum(),
                total_energy=(Ui[4] * dx * dy).sum())


@pytest.mark.parametrize("name,ic,nx,ny,t_end,gamma,lim,bc", [
    ("orszag_tang_fast", "orszag_tang", 32, 32, 0.05, 5 / 3, MC, (PERIODIC, PERIODIC)),
    ("sod_fast", "sod", 64, 1, 0.1, 1.4, MINMOD, (OUTFLOW, PERIODIC)),
    ("brio_wu_fast", "brio_wu", 64, 1, 0.05, 2.0, MINMOD, (OUTFLOW, PERIODIC)),
    ("alfven_fast", "alfven_wave", 32, 1, 1.0, 5 / 3, MC, (PERIODIC, PERIODIC)),
])
def test_verification_ref_scalars(oracle, name, ic, nx, ny, t_end, gamma, lim, bc):
    # include/verification/problems.hpp:28-114: 1-D strips have y_max = 1/nx
    g = make_grid(nx, ny, y_max=1.0 if ny > 1 else 1.0 / nx)
    U0 = oracle.init_ic(ic, g, gamma=gamma, bc=bc)
    U, info = oracle.solve(g, U0, make_params(t_end=t_end, gamma=gamma, limiter=lim), bc=bc)
    ref = read_ref_scalars(GOLD / "verification" / f"{name}.ref")
    d = diagnostics(U, g, gamma)
    assert info["steps"] == int(ref["steps"])
    for key in ("min_rho", "max_rho", "min_pressure", "max_pressure", "mass", "total_energy"):
        assert d[key] == pytest.approx(ref[key], rel=1e-9, abs=1e-12), key


@pytest.mark.slow
def test_orszag_tang_256_published_step_count(oracle):
    # docs/benchmarks.md:83: OT 256^2, MC, t = 0.5 -> 3452 steps
    g = make_grid(256, 256)
    U0 = oracle.init_ic("orszag_tang", g)
    _, info = oracle.solve(g, U0, make_params(t_end=0.5, limiter=MC))
    assert info["steps"] == 3452


# ---- the reference compiled from source -----------------------------------------------------

def _cases(oracle):
    out = []
    g = make_grid(40, 40)
    U = oracle.init_ic("orszag_tang", g)
    m = np.ones((40, 40), np.uint8)
    m[15:25, 12:20] = 0
    oracle.wall_fill(g, U, m, 5 / 3)
    out.append(("mask", g, U, make_params(t_end=0.03), (PERIODIC, PERIODIC), m, True))
    g2 = make_grid(44, 36)
    U2 = oracle.init_ic("rotor", g2, bc=(OUTFLOW, OUTFLOW))
    m2 = np.ones((36, 44), np.uint8)
    m2[0:5, 0:3] = 0
    m2[30:36, 40:44] = 0
    m2[10, 20] = 0
    oracle.wall_fill(g2, U2, m2, 5 / 3)
    out.append(("mask-outflow", g2, U2, make_params(t_end=0.02, limiter=MINMOD),
                (OUTFLOW, OUTFLOW), m2, True))
    out.append(("mask-nowalls", g2, U2, make_params(t_end=0.02), (OUTFLOW, OUTFLOW), m2, False))
    g3 = make_grid(32, 32)
    U3 = oracle.init_ic("orszag_tang", g3)
    out.append(("nonideal", g3, U3, make_params(t_end=0.03, eta_ohm=2e-3, eta_hall=1e-3,
                                                  eta_ambipolar=5e-4),
                (PERIODIC, PERIODIC), None, True))
    g4 = make_grid(64, 1, y_max=1 / 64)
    U4 = oracle.init_ic("brio_wu", g4, gamma=2.0, bc=(OUTFLOW, PERIODIC))
    out.append(("brio_wu", g4, U4, make_params(t_end=0.05, gamma=2.0, limiter=MINMOD),
                (OUTFLOW, PERIODIC), None, True))
    g5 = make_grid(24, 30, x_min=-0.5, x_max=0.7, y_min=0.1, y_max=0.9)
    U5 = oracle.init_ic("blast", g5)
    out.append(("blast-rect", g5, U5, make_params(t_end=0.01, glm_alpha=0.0),
                (PERIODIC, PERIODIC), None, True))
    return out


def test_solve_bitexact_vs_reference(oracle, reference):
    for name, g, U0, p, bc, m, walls in _cases(oracle):
        Uo, io = oracle.solve(g, U0, p, bc=bc, mask=m, walls=walls, dt_cap=4096)
        Ur, ir = reference.solve(g, U0, p, bc=bc, mask=m, walls=walls, dt_cap=4096)
        assert io["steps"] == ir["steps"], name
        assert io["c_h"] == ir["c_h"], name
        np.testing.assert_array_equal(io["dt"], ir["dt"], err_msg=name)
        assert np.array_equal(Uo, Ur), name


def test_rhs_cfl_bitexact_vs_reference(oracle, reference):
    for name, g, U0, p, bc, m, walls in _cases(oracle):
        for ch in (0.0, 1.7):
            assert np.array_equal(oracle.rhs(g, U0, p, ch, m), reference.rhs(g, U0, p, ch, m)), name
        assert oracle.cfl(g, U0, p, m) == reference.cfl(g, U0, p, m), name


def test_initial_conditions_bitexact_vs_reference(oracle, reference):
    names = ["uniform", "sine", "blast", "orszag_tang", "sod", "brio_wu", "rotor",
             "kelvin_helmholtz", "alfven_wave"]
    g = make_grid(37, 29, x_min=-0.3, x_max=1.1, y_min=0.05, y_max=0.95)
    for n in names:
        Uo = oracle.init_ic(n, g, bc=(PERIODIC, PERIODIC))
        Ur = reference.init_ic(n, g)
        reference.apply_bc(g, Ur, (PERIODIC, PERIODIC))
        assert np.array_equal(Uo, Ur), n


def test_face_flux_random_states_bitexact(oracle, reference):
    rng = np.random.default_rng(7)
    for k in range(400):
        W = rng.normal(size=(4, 9))
        W[:, 0] = np.abs(W[:, 0]) + 0.05
        W[:, 4] = np.abs(W[:, 4]) + 0.05
        if k % 7 == 0:
            W[1, 0] = -0.2   # forces the first-order fallback
        for lim in (MINMOD, MC):
            a = oracle.face_flux(lim, *W, 5 / 3, 0.8)
            b = reference.face_flux(lim, *W, 5 / 3, 0.8)
            assert np.array_equal(a, b)


def test_wall_fill_bitexact_vs_reference(oracle, reference):
    rng = np.random.default_rng(3)
    g = make_grid(23, 17)
    U = oracle.init_ic("rotor", g)
    for _ in range(5):
        m = (rng.random((17, 23)) > 0.35).astype(np.uint8)
        a = U.copy()
        b = U.copy()
        oracle.wall_fill(g, a, m, 1.4)
        reference.wall_fill(g, b, m, 1.4)
        assert np.array_equal(a, b)


# ---- mathematics ---------------------------------------------------------------------------

def test_limiter_closed_forms(oracle):
    lim = oracle.lib.orc_limit
    assert lim(MINMOD, 1.0, 2.0) == 1.0
    assert lim(MINMOD, -3.0, -2.0) == -2.0
    assert lim(MINMOD, 1.0, -1.0) == 0.0
    assert lim(MINMOD, 0.0, 5.0) == 0.0
    assert lim(MINMOD, 2.0, 2.0) == 2.0
    assert lim(MINMOD, 1e-300, 1e-300) == 0.0      # a*b underflows to 0 -> flat
    # MC = minmod((a+b)/2, minmod(2a, 2b))
    assert lim(MC, 1.0, 2.0) == 1.5
    assert lim(MC, 1.0, 10.0) == 2.0
    assert lim(MC, -1.0, -10.0) == -2.0
    assert lim(MC, 1.0, -3.0) == 0.0


def test_hll_consistency_and_glm_terms(oracle):
    # equal states: the HLL flux equals the physical GLM flux (include/mhd/mhd_physics.hpp:62-85)
    rng = np.random.default_rng(11)
    for _ in range(50):
        W = rng.normal(size=9)
        W[0] = abs(W[0]) + 0.1
        W[4] = abs(W[4]) + 0.1
        gam, ch = 1.4, 0.9
        F = oracle.face_flux(MC, W, W, W, W, gam, ch)
        rho, vn, vt, vz, p, bn, bt, bz, psi = W
        pm = 0.5 * (bn * bn + bt * bt + bz * bz)
        E = p / (gam - 1) + 0.5 * rho * (vn * vn + vt * vt + vz * vz) + pm
        vb = vn * bn + vt * bt + vz * bz
        exact = [rho * vn, rho * vn * vn + p + pm - bn * bn, rho * vn * vt - bn * bt,
                 rho * vn * vz - bn * bz, (E + p + pm) * vn - bn * vb + psi * bn, psi,
                 vn * bt - vt * bn, vn * bz - vz * bn, ch * ch * bn]
        np.testing.assert_allclose(F, exact, rtol=1e-12, atol=1e-12)


def test_uniform_state_invariance_and_conservation(oracle):
    g = make_grid(16, 12)
    U = empty_state(g)
    U[0] = 1.3
    U[1] = 0.2
    U[2] = -0.1
    U[5] = 0.4
    U[6] = 0.3
    U[4] = 2.0 / (5 / 3 - 1) + 0.5 * (0.2 ** 2 + 0.1 ** 2) / 1.3 + 0.5 * (0.4 ** 2 + 0.3 ** 2)
    R = oracle.rhs(g, U, make_params(), 1.2)
    assert np.abs(R).max() < 1e-13
    g = make_grid(48, 40)
    U0 = oracle.init_ic("orszag_tang", g)
    U, _ = oracle.solve(g, U0, make_params(t_end=0.05, glm_alpha=0.0))
    for c in range(8):
        a, b = interior(U0)[c].sum(), interior(U)[c].sum()
        assert abs(a - b) <= 1e-11 * max(1.0, np.abs(interior(U0)[c]).sum()), FIELDS[c]


def test_nonideal_field_closed_form(oracle):
    J = np.array([0.3, -1.2, 0.7])
    B = np.array([0.5, 0.25, -0.8])
    p = make_params(eta_ohm=0.1, eta_hall=0.05, eta_ambipolar=0.02)
    E = np.zeros(3)
    oracle.lib.orc_nonideal_e(J, B, p, E)
    b = np.linalg.norm(B)
    jxb = np.cross(J, B)
    exact = 0.1 * J + 0.05 * jxb / b + 0.02 * np.cross(jxb, B) / b ** 2
    np.testing.assert_allclose(E, exact, rtol=1e-13)
    E2 = np.zeros(3)
    oracle.lib.orc_nonideal_e(J, np.array([1e-8, 0.0, 0.0]), p, E2)   # |B|^2 < 1e-14
    np.testing.assert_allclose(E2, 0.1 * J, rtol=1e-15)


def test_ghost_indexing_integer_tags(oracle):
    # every cell holds its own index; after the fill a ghost must hold its source's tag
    for nx, ny, bcx, bcy in [(5, 4, PERIODIC, PERIODIC), (6, 1, OUTFLOW, PERIODIC),
                             (1, 3, PERIODIC, OUTFLOW), (7, 5, OUTFLOW, OUTFLOW)]:
        g = make_grid(nx, ny)
        U = empty_state(g)
        nxt, nyt = nx + 4, ny + 4
        tags = np.arange(nxt * nyt, dtype=np.float64).reshape(nyt, nxt)
        U[:] = tags
        oracle.apply_bc(g, U, (bcx, bcy))

        def src(k, n, bc):
            if 2 <= k < 2 + n:
                return k
            if bc == PERIODIC:
                return 2 + ((k - 2) % n)
            return 2 if k < 2 else 1 + n

        for j in range(nyt):
            for i in range(nxt):
                expect = tags[src(j, ny, bcy), src(i, nx, bcx)]
                assert U[0, j, i] == expect, (nx, ny, i, j)


def test_periodic_source_small_n(oracle):
    # include/boundary_conditions/boundary_conditions.hpp:26-30, tests/test_boundary_conditions.cpp:73-76
    ps = oracle.lib.orc_periodic_source
    assert [ps(g, 1, 2, 1) for g in range(2)] == [2, 2]
    assert [ps(g, 1, 2, 0) for g in range(2)] == [2, 2]
    assert [ps(g, 5, 2, 1) for g in range(2)] == [5, 6]
    assert [ps(g, 5, 2, 0) for g in range(2)] == [2, 3]


# ---- rare branches of the face kernel on whole solves ---------------------------------------

def test_fallback_and_degenerate_span_solves_bitexact_vs_reference(oracle, reference):
    """The states that force the positivity fallback (a block of p < 0) and the degenerate
    wave-speed span (a cold flow with c_h < 5e-15) solve for a few steps to the reference's
    bits, with finite states; c_h of the cold flow stays below the 1e-14 span bound."""
    from rare_states import cold_slow_flow, negative_pressure_patch
    g = make_grid(48, 40)
    for name, U0, lim in (("fallback", negative_pressure_patch(oracle, g), MC),
                          ("degenerate", cold_slow_flow(oracle, g), MINMOD)):
        p = make_params(t_end=1e300, limiter=lim, max_steps=4)
        a, ia = oracle.solve(g, U0, p, allow_max_steps=True, dt_cap=8)
        b, ib = reference.solve(g, U0, p, allow_max_steps=True, dt_cap=8)
        assert np.isfinite(a).all(), name
        assert ia["steps"] == ib["steps"] == 4 and ia["c_h"] == ib["c_h"], name
        np.testing.assert_array_equal(ia["dt"], ib["dt"])
        assert np.array_equal(a, b), name
        if name == "degenerate":
            assert 0.0 < ia["c_h"] < 5e-15
            assert not np.array_equal(a, U0)   # the flow moved


def test_degenerate_span_average_matches_closed_form(oracle):
    """A face whose speeds span less than 1e-14 (sL < 0 < sR) returns (FL + FR) / 2 exactly,
    where the HLL formula would add the dissipation sL sR (UR - UL) / (sR - sL)."""
    W = np.zeros((4, 9))
    W[:, 0] = [1.0, 1.1, 0.9, 1.0]
    W[:, 4] = [1e-32, 2e-32, 1.5e-32, 1e-32]
    W[:, 1] = [1e-18, -1e-18, 2e-18, 0.0]
    F = oracle.face_flux(MINMOD, *W, 5.0 / 3.0, 1e-16)
    # minmod reconstruction of the face between cells 1 and 2
    def lim(a, b):
        return 0.0 if a * b <= 0 else (a if abs(a) < abs(b) else b)
    WL = W[1] + 0.5 * np.array([lim(W[1, k] - W[0, k], W[2, k] - W[1, k]) for k in range(9)])
    WR = W[2] - 0.5 * np.array([lim(W[2, k] - W[1, k], W[3, k] - W[2, k]) for k in range(9)])

    def flux(w):
        rho, vn, p = w[0], w[1], w[4]
        e = p / (5.0 / 3.0 - 1.0) + 0.5 * rho * vn * vn
        return np.array([rho * vn, rho * vn * vn + p, 0, 0, (e + p) * vn, 0, 0, 0, 0])
    np.testing.assert_allclose(F, 0.5 * (flux(WL) + flux(WR)), rtol=1e-14, atol=1e-60)


def test_forced_dt_solve_reproduces_the_cfl_solve(oracle):
    """orc_solve_dt driven by a solve's own dt sequence reproduces that solve bit for bit (the
    mode the contracted fp64 build is compared in), and a different dt gives another state."""
    g = make_grid(40, 36)
    U0 = oracle.init_ic("orszag_tang", g)
    p = make_params(t_end=0.03, limiter=MC, eta_ohm=1e-3, eta_hall=1e-3)
    U1, info = oracle.solve(g, U0, p, dt_cap=64)
    U2, info2 = oracle.solve_dt(g, U0, p, info["dt"])
    assert info2["steps"] == info["steps"] and info2["t"] == info["t"]
    assert info2["c_h"] == info["c_h"]
    assert np.array_equal(U1, U2)
    U3, _ = oracle.solve_dt(g, U0, p, info["dt"] * (1.0 + 1e-9))
    assert not np.array_equal(U1, U3)
