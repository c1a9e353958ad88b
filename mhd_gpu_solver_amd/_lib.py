"""ctypes binding of libmhd_hip.so (the C ABI declared in include/mhd_hip.h).

The library is loaded from this package directory only. There is no fallback: importing a
name that needs the library raises a RuntimeError when it has not been built.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import threading

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libmhd_hip.so"

OK, EINVAL, ERUNTIME, ENODEV, EFULL = 0, 1, 2, 3, 4
FP64, FP32, FP64_FAST = 0, 1, 2
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp64_fast": FP64_FAST}
FILL_GHOSTS, FILL_WALLS = 1, 2


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("ng", C.c_int), ("x_min", C.c_double),
                ("x_max", C.c_double), ("y_min", C.c_double), ("y_max", C.c_double)]


class Params(C.Structure):
    _fields_ = [("t_end", C.c_double), ("cfl", C.c_double), ("gamma", C.c_double),
                ("glm_alpha", C.c_double), ("limiter", C.c_int), ("ohmic", C.c_int),
                ("hall", C.c_int), ("ambipolar", C.c_int), ("eta_ohm", C.c_double),
                ("eta_hall", C.c_double), ("eta_ambipolar", C.c_double),
                ("max_steps", C.c_int)]


class Result(C.Structure):
    _fields_ = [("t", C.c_double), ("steps", C.c_int), ("c_h", C.c_double)]


class Instance(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("x_min", C.c_double), ("x_max", C.c_double),
                ("y_min", C.c_double), ("y_max", C.c_double), ("bc_x", C.c_int),
                ("bc_y", C.c_int), ("t_end", C.c_double), ("cfl", C.c_double),
                ("gamma", C.c_double), ("glm_alpha", C.c_double), ("max_steps", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("instances", C.c_int), ("active", C.c_int), ("tiles", C.c_int),
                ("seg_rows", C.c_int), ("blocks", C.c_int), ("bytes_used", C.c_size_t),
                ("bytes_budget", C.c_size_t), ("last_seconds", C.c_double),
                ("paused", C.c_int)]


DP9 = C.POINTER(C.c_double) * 9

_lock = threading.Lock()
_lib = None

# (name, restype, argtypes) of every entry point declared in include/mhd_hip.h
SIGNATURES = [
    ("mhd_hip_available", C.c_int, []),
    ("mhd_hip_last_error", C.c_char_p, []),
    ("mhd_hip_create", C.c_int, [C.POINTER(Grid), C.c_int, C.POINTER(C.c_void_p)]),
    ("mhd_hip_destroy", None, [C.c_void_p]),
    ("mhd_hip_device_name", C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    ("mhd_hip_precision", C.c_int, [C.c_void_p]),
    ("mhd_hip_configure", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_uint8),
                                    C.c_size_t, C.c_int]),
    ("mhd_hip_upload", C.c_int, [C.c_void_p, DP9, C.c_size_t]),
    ("mhd_hip_download", C.c_int, [C.c_void_p, DP9, C.c_size_t]),
    ("mhd_hip_solve", C.c_int, [C.c_void_p, C.POINTER(Params), C.POINTER(Result)]),
    ("mhd_hip_advance_steps", C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int,
                                        C.POINTER(Result)]),
    ("mhd_hip_set_dt_log", C.c_int, [C.c_void_p, C.c_int]),
    ("mhd_hip_dt_log", C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int,
                                 C.POINTER(C.c_int)]),
    ("mhd_hip_fill_boundaries", C.c_int, [C.c_void_p, C.c_double, C.c_int]),
    ("mhd_hip_last_solve_seconds", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("mhd_hip_totals", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("mhd_hip_totals_noalloc", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("mhd_hip_set_seg_rows", C.c_int, [C.c_void_p, C.c_int]),
    ("mhd_hip_get_seg_rows", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("mhd_hip_set_kernel_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("mhd_hip_kernel_times", C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                       C.POINTER(C.c_longlong)]),
    ("mhd_hip_init_ic", C.c_int, [C.c_char_p, C.POINTER(Grid), C.c_double, C.c_double,
                                  C.c_uint64, DP9]),
    ("mhd_hip_wall_plan", C.c_int, [C.POINTER(Grid), C.POINTER(C.c_uint8), C.c_size_t,
                                    C.POINTER(C.c_int64), C.c_size_t, C.POINTER(C.c_size_t)]),
    ("mhd_hip_batch_instance_bytes", C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    ("mhd_hip_batch_create", C.c_int, [C.c_int, C.POINTER(Params), C.c_size_t, C.c_int,
                                       C.POINTER(C.c_void_p)]),
    ("mhd_hip_batch_destroy", None, [C.c_void_p]),
    ("mhd_hip_batch_add", C.c_int, [C.c_void_p, C.POINTER(Instance), DP9, C.POINTER(C.c_int)]),
    ("mhd_hip_batch_add_config", C.c_int, [C.c_void_p, C.POINTER(Instance), C.POINTER(Params),
                                           C.POINTER(C.c_uint8), C.c_size_t, C.c_int, DP9,
                                           C.POINTER(C.c_int)]),
    ("mhd_hip_batch_remove", C.c_int, [C.c_void_p, C.c_int]),
    ("mhd_hip_batch_advance", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_int)]),
    ("mhd_hip_batch_prepare", C.c_int, [C.c_void_p]),
    ("mhd_hip_batch_result", C.c_int, [C.c_void_p, C.c_int, C.POINTER(Result),
                                       C.POINTER(C.c_int)]),
    ("mhd_hip_batch_download", C.c_int, [C.c_void_p, C.c_int, DP9, C.c_size_t]),
    ("mhd_hip_batch_totals", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    ("mhd_hip_batch_set_seg_rows", C.c_int, [C.c_void_p, C.c_int]),
    ("mhd_hip_batch_info", C.c_int, [C.c_void_p, C.POINTER(BatchStats)]),
    ("mhd_hip_batch_max_div_b", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    ("mhd_hip_max_div_b", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("mhd_hip_stream_bandwidth", C.c_int, [C.c_int, C.c_longlong, C.c_int, C.c_int,
                                           C.POINTER(C.c_double)]),
    ("mhd_hip_rel_l1", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    ("mhd_hip_batch_rel_l1", C.c_int,
     [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    ("mhd_hip_batch_set_paused", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("mhd_hip_batch_progress", C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                         C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    ("mhd_hip_batch_set_schedule", C.c_int, [C.c_void_p, C.c_int, C.c_longlong]),
    ("mhd_hip_batch_poll_finished", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int,
                                              C.POINTER(C.c_int)]),
    ("mhd_hip_batch_set_kernel_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("mhd_hip_batch_kernel_times", C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                             C.POINTER(C.c_longlong)]),
]


class CapacityFull(RuntimeError):
    """MHD_HIP_EFULL: the batch's HBM budget or instance slots are exhausted; retry later."""


def lib() -> C.CDLL:
    """The loaded library; raises RuntimeError if it was not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build the hip backend with "
                    "`python -m mhd_gpu_solver_amd.build` (hipcc, gfx950)")
            handle = C.CDLL(str(LIB_PATH))
            for name, res, args in SIGNATURES:
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
        return _lib


def last_error() -> str:
    return lib().mhd_hip_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    """Map a status code to the exception type the reference raises."""
    if rc == OK:
        return
    msg = last_error() or what
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == EFULL:
        raise CapacityFull(msg)
    raise RuntimeError(msg)


def planes(arr, writable=False):
    """Nine plane pointers into a C-contiguous (9, ny_tot, nx_tot) float64 array."""
    if arr.dtype.name != "float64" or not arr.flags.c_contiguous or arr.shape[0] != 9:
        raise ValueError("state must be a C-contiguous float64 array of shape (9, ny, nx)")
    if writable and not arr.flags.writeable:
        raise ValueError("state array is read-only")
    n = arr.shape[1] * arr.shape[2]
    base = arr.ctypes.data
    ptrs = DP9(*[C.cast(base + c * n * 8, C.POINTER(C.c_double)) for c in range(9)])
    return ptrs, n
