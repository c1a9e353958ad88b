"""The reference's GPU acceptance gate and backend-contract tests, restated on the hip backend.

Gate (tools/mhd_backend_compare.cpp:242-250, 301-416; docs/gpu.md:125-137), per precision:
* Orszag-Tang 128^2 (the reference IC, B0 = 1, MC, CFL 0.4, alpha 0.1, periodic) to t = 0.25,
  ideal and Ohmic 1e-3 + Hall 5e-4: the same step count as the fp64 CPU path (the oracle);
  relative L1 of rho, p and (bx, by) against it <= 1e-11 (fp64) / 1e-5 (fp32); relative L-inf
  of rho <= 1e-9 / 5e-5; mass and energy drift <= 1e-11 / 2e-6.
* Alfven wave (delta = 1e-3, ny = 1 strip, MC) to t = 1: observed order N = 128 -> 256 >= 1.8,
  and the N = 256 error over the CPU path's error <= 1.001 (fp64) / 1.25 (fp32).
Norms follow the reference's compare_states / alfven_wave_l1_error: primitives by
to_primitive (momentum divided by rho), interior cells only.

Contract (tests/test_gpu_backends.cpp:126-157): RepeatedRunsAccumulateTime and
MaxStepsIsEnforced, through MHDSolver and through the C ABI's status codes.
"""
import ctypes as C
import math

import numpy as np
import pytest

from checkers import MC, make_grid, make_params

pytestmark = pytest.mark.gpu

GATE = {"fp64": {"l1": 1e-11, "linf": 1e-9, "drift": 1e-11, "alfven_ratio": 1.001},
        "fp32": {"l1": 1e-5, "linf": 5e-5, "drift": 2e-6, "alfven_ratio": 1.25}}


def primitives(U, gamma):
    """rho, p, bx, by of the interior (include/mhd/mhd_physics.hpp:23-39 order)."""
    u = U[:, 2:-2, 2:-2]
    rho = u[0]
    vx, vy, vz = u[1] / rho, u[2] / rho, u[3] / rho
    e_kin = 0.5 * rho * (vx * vx + vy * vy + vz * vz)
    e_mag = 0.5 * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7])
    p = (gamma - 1.0) * (u[4] - e_kin - e_mag)
    return rho, p, u[5], u[6]


def compare_states(a, ref, gamma):
    """tools/mhd_backend_compare.cpp:169-196"""
    ra, pa, bxa, bya = primitives(a, gamma)
    rr, pr, bxr, byr = primitives(ref, gamma)
    rho_l1 = np.abs(ra - rr).sum() / np.abs(rr).sum()
    p_l1 = np.abs(pa - pr).sum() / np.abs(pr).sum()
    b_l1 = (np.abs(bxa - bxr).sum() + np.abs(bya - byr).sum()) / \
        (np.abs(bxr).sum() + np.abs(byr).sum())
    rho_linf = np.abs(ra - rr).max() / np.abs(rr).max()
    return max(rho_l1, p_l1, b_l1), rho_linf


def setup_orszag(s, nonideal):
    """setup_orszag / setup_orszag_nonideal (tools/mhd_backend_compare.cpp:219-234)"""
    s.set_cfl(0.4)
    s.set_glm_alpha(0.1)
    s.set_limiter_mc()
    s.set_periodic_bc()
    s.set_ideal()
    s.initialize("orszag_tang")
    if nonideal:
        s.enable_ohmic(1e-3)
        s.enable_hall(5e-4)


@pytest.mark.parametrize("nonideal", [False, True], ids=["ideal", "ohmic_hall"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gate_orszag_tang_128(hip, oracle, precision, nonideal):
    n, t_acc = 128, 0.25
    s = hip.MHDSolver(n, n)
    setup_orszag(s, nonideal)
    s.set_backend("hip", precision)
    U0 = s.state.copy()
    r = s.run(t_acc)
    p = make_params(t_end=t_acc, cfl=0.4, gamma=5.0 / 3.0, glm_alpha=0.1, limiter=MC,
                    eta_ohm=1e-3 if nonideal else None, eta_hall=5e-4 if nonideal else None)
    Uo, info = oracle.solve(make_grid(n, n), U0, p)
    th = GATE[precision]
    l1, linf = compare_states(s.state, Uo, 5.0 / 3.0)
    i = (slice(2, -2), slice(2, -2))
    mass0, energy0 = U0[0][i].sum(), U0[4][i].sum()
    dm = abs(s.state[0][i].sum() - mass0) / mass0
    de = abs(s.state[4][i].sum() - energy0) / energy0
    print(f"gate {precision} OT128 {'ohmic+hall' if nonideal else 'ideal'}: steps {r.steps} "
          f"vs {info['steps']}, L1 {l1:.2e}, Linf {linf:.2e}, drift {dm:.1e} / {de:.1e}")
    assert r.steps == info["steps"]
    assert l1 <= th["l1"], l1
    assert linf <= th["linf"], linf
    assert max(dm, de) <= th["drift"], (dm, de)


def alfven_error(U, n, t, gamma=5.0 / 3.0):
    """include/verification/diagnostics.hpp:87-107 against the exact wave
    (include/initial_condition/inititial_condition.hpp:175-192)."""
    u = U[:, 2:-2, 2:-2]
    rho = u[0]
    x = (np.arange(n) + 0.5) / n
    ph = 2.0 * math.pi * (x - t)
    d = 1e-3
    exact = [d * np.cos(ph), d * np.sin(ph), -d * np.cos(ph), -d * np.sin(ph)]
    got = [u[6], u[7], u[2] / rho, u[3] / rho]
    num = sum(np.abs(g - e[None, :]).sum() for g, e in zip(got, exact))
    den = sum(np.abs(e[None, :]).sum() * u.shape[1] for e in exact)
    return num / den


def alfven_run(hip, oracle, n, precision):
    """make_fast_alfven (include/verification/problems.hpp:100-114) at N, MC, ideal, t = 1."""
    s = hip.MHDSolver(n, 1, 0.0, 1.0, 0.0, 1.0 / n)
    s.set_cfl(0.4)
    s.set_gamma(5.0 / 3.0)
    s.set_glm_alpha(0.1)
    s.set_periodic_bc()
    s.set_limiter_mc()
    s.set_ideal()
    s.initialize("alfven_wave")
    U0 = s.state.copy()
    if precision is None:  # the CPU path
        Uo, info = oracle.solve(make_grid(n, 1, 0.0, 1.0, 0.0, 1.0 / n), U0,
                                make_params(t_end=1.0, limiter=MC))
        return alfven_error(Uo, n, info["t"])
    s.set_backend("hip", precision)
    r = s.run(1.0)
    return alfven_error(s.state, n, r.t)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gate_alfven_order_and_error_ratio(hip, oracle, precision):
    e128 = alfven_run(hip, oracle, 128, precision)
    e256 = alfven_run(hip, oracle, 256, precision)
    cpu256 = alfven_run(hip, oracle, 256, None)
    order = math.log(e128 / e256) / math.log(2.0)
    ratio = e256 / cpu256
    print(f"gate {precision} Alfven: e128 {e128:.3e} e256 {e256:.3e} (CPU {cpu256:.3e}), "
          f"order {order:.3f}, ratio {ratio:.4f}")
    assert order >= 1.8
    assert ratio <= GATE[precision]["alfven_ratio"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_repeated_runs_accumulate_time(hip, precision):
    """RepeatedRunsAccumulateTime (tests/test_gpu_backends.cpp:126-141)"""
    s = hip.MHDSolver(32, 32)
    s.initialize("orszag_tang")
    s.set_backend("hip", precision)
    r1 = s.run(0.01)
    r = s.advance_to(0.02)
    assert r.t == pytest.approx(0.02, abs=1e-12)
    assert r.steps > r1.steps > 0
    with pytest.raises(ValueError):
        s.run(-1.0)
    with pytest.raises(ValueError):
        s.advance_to(0.01)
    assert s.time == r.t and s.steps == r.steps


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_max_steps_is_enforced(hip, precision):
    """MaxStepsIsEnforced (tests/test_gpu_backends.cpp:143-157): the solve stops at max_steps
    before t_end with a runtime error, and the host state and clock stay where they were."""
    s = hip.MHDSolver(32, 32)
    s.initialize("orszag_tang")
    s.params.max_steps = 3
    s.set_backend("hip", precision)
    U0 = s.state.copy()
    with pytest.raises(RuntimeError):
        s.run(1.0)
    assert s.time == 0.0 and s.steps == 0
    assert np.array_equal(s.state, U0)
    s.params.max_steps = 1_000_000
    assert s.run(0.005).steps > 0   # the solver stays usable after the error


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_solve_status_codes_through_the_c_abi(hip, precision):
    """mhd_hip_solve maps the reference's exceptions to status codes: t_end < 0 and cfl <= 0
    -> EINVAL (std::invalid_argument), max_steps before t_end -> ERUNTIME (std::runtime_error,
    device_solver.hpp:57-75, cuda_solver.cu:601-608,647-649)."""
    from mhd_gpu_solver_amd import _lib
    lib = _lib.lib()
    n = 32
    g = _lib.Grid(n, n, 2, 0.0, 1.0, 0.0, 1.0)
    h = C.c_void_p()
    assert lib.mhd_hip_create(C.byref(g), _lib.FP64 if precision == "fp64" else _lib.FP32,
                              C.byref(h)) == _lib.OK
    try:
        U = hip.init_ic("orszag_tang", n, n)
        assert lib.mhd_hip_configure(h, 0, 0, None, 0, 1) == _lib.OK
        ptrs, cnt = _lib.planes(U)
        assert lib.mhd_hip_upload(h, ptrs, cnt) == _lib.OK
        res = _lib.Result()
        p = hip.SolveParams(t_end=-1.0).to_c()
        assert lib.mhd_hip_solve(h, C.byref(p), C.byref(res)) == _lib.EINVAL
        assert "t_end" in _lib.last_error()
        p = hip.SolveParams(t_end=0.01, cfl=0.0).to_c()
        assert lib.mhd_hip_solve(h, C.byref(p), C.byref(res)) == _lib.EINVAL
        p = hip.SolveParams(t_end=1.0, max_steps=3).to_c()
        assert lib.mhd_hip_solve(h, C.byref(p), C.byref(res)) == _lib.ERUNTIME
        assert "max_steps" in _lib.last_error()
        p = hip.SolveParams(t_end=0.01).to_c()
        assert lib.mhd_hip_solve(h, C.byref(p), C.byref(res)) == _lib.OK
        assert res.steps > 0 and res.t == pytest.approx(0.01, abs=1e-15)
    finally:
        lib.mhd_hip_destroy(h)
