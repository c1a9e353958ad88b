"""CPU-side checks of the hip backend: the C-ABI library loads and exports every entry point of
include/mhd_hip.h, and the host logic that needs no device (initial conditions, wall-fill plan,
argument validation) matches the oracle bit for bit."""
import ctypes as C
import pathlib
import re

import numpy as np
import pytest

import mhd_gpu_solver_amd as m
from mhd_gpu_solver_amd import _lib
from checkers import interior, make_grid

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(mhd_hip_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    decl = declared_symbols()
    assert len(decl) >= 15
    for name in sorted(decl):
        assert hasattr(lib, name), name
    bound = {s[0] for s in _lib.SIGNATURES}
    assert decl == bound, decl ^ bound


@pytest.mark.parametrize("name", ["uniform", "sine", "blast", "orszag_tang", "sod", "brio_wu",
                                  "rotor", "kelvin_helmholtz", "alfven_wave", "kh", "alfven"])
def test_reference_ics_match_oracle(oracle, name):
    g = make_grid(37, 29, x_min=-0.3, x_max=1.1, y_min=0.05, y_max=0.95)
    U = m.init_ic(name, 37, 29, -0.3, 1.1, 0.05, 0.95, gamma=1.4)
    Uo = oracle.init_ic(name, g, gamma=1.4)
    assert np.array_equal(interior(U), interior(Uo))


def test_parametrised_ics():
    b0 = 1.0 / np.sqrt(4 * np.pi)
    U1 = m.init_ic("orszag_tang", 32, 32)
    Ub = m.init_ic("orszag_tang_b0", 32, 32, b0=1.0)
    assert np.array_equal(U1, Ub)        # b0 = 1 is the reference IC
    Us = m.init_ic("orszag_tang_b0", 32, 32, b0=b0)
    np.testing.assert_allclose(Us[5], b0 * U1[5], rtol=1e-15)
    R1 = m.init_ic("rotor", 32, 32)
    Rb = m.init_ic("rotor_b0", 32, 32, b0=1.0)
    assert np.array_equal(R1, Rb)
    K1 = m.init_ic("kh_tanh", 32, 32, seed=5)
    K2 = m.init_ic("kh_tanh", 32, 32, seed=5)
    K3 = m.init_ic("kh_tanh", 32, 32, seed=6)
    assert np.array_equal(K1, K2) and not np.array_equal(K1, K3)
    vy = interior(K1)[2] / interior(K1)[0]
    assert np.abs(vy).max() <= 1e-2 and np.abs(vy).max() > 5e-3
    A = m.init_ic("alfven_oblique", 64, 64)
    rho = interior(A)[0]
    assert np.allclose(rho, 1.0)
    # |B|^2 = 1 + amp^2 everywhere for the circularly polarised wave
    b2 = (interior(A)[5:8] ** 2).sum(axis=0)
    np.testing.assert_allclose(b2, 1.01, rtol=1e-13)


def test_unknown_ic_raises():
    with pytest.raises(ValueError):
        m.init_ic("no_such_ic", 8, 8)


def test_wall_plan_matches_oracle_fill(oracle):
    rng = np.random.default_rng(5)
    for trial in range(6):
        nx, ny = 19 + trial, 13
        mask = (rng.random((ny, nx)) > 0.4).astype(np.uint8)
        plan = m.wall_plan(mask, nx, ny)
        assert len(plan) == int((mask == 0).sum())
        # apply the plan with tag arithmetic: dst must take src, as the oracle's fill chooses
        g = make_grid(nx, ny)
        U = oracle.init_ic("rotor", g)
        # mark each cell with a unique density so the chosen source is identifiable
        nxt = nx + 4
        U[0] = 1.0 + np.arange(U[0].size).reshape(U[0].shape) * 1e-3
        U[1:4] = 0.0
        F = U.copy()
        oracle.wall_fill(g, F, mask, 5 / 3)
        for dst, src, axis in plan:
            j, i = divmod(int(dst), nxt)
            if src < 0:
                assert F[0, j, i] == 1.0
            else:
                sj, si = divmod(int(src), nxt)
                assert F[0, j, i] == U[0, sj, si]


def test_wall_plan_size_mismatch():
    with pytest.raises(ValueError):
        m.wall_plan(np.ones((3, 4), np.uint8).ravel()[:5], 4, 3)


def test_solver_argument_validation_without_device():
    with pytest.raises(ValueError):
        m.MHDSolver(8, 8, ng=1)
    s = m.MHDSolver(8, 6)
    with pytest.raises(ValueError):
        s.set_mask(np.ones((8, 6), np.uint8))
    with pytest.raises(ValueError):
        s.set_field("rho", np.ones((3, 3)))
    with pytest.raises(ValueError):
        s.field("temperature")
    with pytest.raises(ValueError):
        s.set_backend("tpu")
    with pytest.raises(RuntimeError):
        s.set_backend("cpu")
    with pytest.raises(ValueError):
        s.run(-1.0)


def test_null_and_bad_arguments_return_codes():
    lib = _lib.lib()
    h = C.c_void_p()
    g = _lib.Grid(8, 8, 1, 0.0, 1.0, 0.0, 1.0)
    assert lib.mhd_hip_create(C.byref(g), 0, C.byref(h)) in (_lib.EINVAL, _lib.ENODEV)
    assert lib.mhd_hip_solve(None, None, None) == _lib.EINVAL
    assert lib.mhd_hip_last_error()


def test_server_rejects_unknown_policy():
    from mhd_gpu_solver_amd.serve import Server
    with pytest.raises(ValueError):
        Server(policy="lifo")


def test_no_cpu_execution_settings():
    """The reference module's CPU thread-pool and SIMD settings (python/mhd_pybind.cpp:169-180)
    steer a CPU solver this package does not have, so they are not offered."""
    for name in ("set_num_threads", "num_threads", "set_simd_enabled", "simd_enabled",
                 "simd_backend", "simd_width"):
        assert not hasattr(m, name), name


def test_stream_bandwidth_argument_validation_without_device():
    """mhd_hip_stream_bandwidth checks its arguments before it touches a device."""
    lib = _lib.lib()
    v = C.c_double(-1.0)
    assert lib.mhd_hip_stream_bandwidth(_lib.FP32, 1000, 3, 5, C.byref(v)) == _lib.EINVAL
    assert lib.mhd_hip_stream_bandwidth(_lib.FP32, 0, 2, 5, C.byref(v)) == _lib.EINVAL
    assert lib.mhd_hip_stream_bandwidth(_lib.FP64, 1000, 1, 0, C.byref(v)) == _lib.EINVAL
    assert lib.mhd_hip_stream_bandwidth(7, 1000, 1, 5, C.byref(v)) == _lib.EINVAL
    assert lib.mhd_hip_stream_bandwidth(_lib.FP64, 1000, 1, 5, None) == _lib.EINVAL
    assert v.value == -1.0
