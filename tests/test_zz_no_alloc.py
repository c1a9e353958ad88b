"""mhd_hip_totals_noalloc (the totals entry point that allocates nothing) through the C ABI:
on two calls in a row, the second on a smaller grid in the other precision, it returns the
same bits as mhd_hip_totals on the same device state (same kernels, same order)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(hip, nx, ny, precision):
    from mhd_gpu_solver_amd import _lib
    lib = _lib.lib()
    g = _lib.Grid(nx, ny, 2, 0.0, 1.0, 0.0, 1.0)
    h = C.c_void_p()
    assert lib.mhd_hip_create(C.byref(g), _lib.FP64 if precision == "fp64" else _lib.FP32,
                              C.byref(h)) == _lib.OK, _lib.last_error()
    U = hip.init_ic("rotor", nx, ny)
    assert lib.mhd_hip_configure(h, 0, 0, None, 0, 1) == _lib.OK
    ptrs, n = _lib.planes(U)
    assert lib.mhd_hip_upload(h, ptrs, n) == _lib.OK
    return h


def _totals(fn, h):
    from mhd_gpu_solver_amd import _lib
    out = (C.c_double * 9)()
    rc = fn(h, out)
    assert rc == _lib.OK, _lib.last_error()
    return np.array(out[:])


def test_totals_noalloc_equals_totals(hip):
    from mhd_gpu_solver_amd import _lib
    lib = _lib.lib()
    cases = [(300, 211, "fp64"), (40, 29, "fp32")]
    handles = []
    try:
        for nx, ny, prec in cases:
            h = _solver(hip, nx, ny, prec)
            handles.append(h)
            # the no-alloc form first, on a solver whose own scratch was never allocated
            a = _totals(lib.mhd_hip_totals_noalloc, h)
            b = _totals(lib.mhd_hip_totals, h)
            c = _totals(lib.mhd_hip_totals_noalloc, h)
            assert np.array_equal(a, b), (nx, ny, prec, a - b)
            assert np.array_equal(c, b)
            assert a[0] > 0 and np.isfinite(a).all()
        assert lib.mhd_hip_totals_noalloc(None, (C.c_double * 9)()) == _lib.EINVAL
    finally:
        for h in handles:
            lib.mhd_hip_destroy(h)
