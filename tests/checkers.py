"""ctypes access to the CPU checkers (test infrastructure only).

* ``Oracle`` wraps ``oracle/libmhd_oracle.so``, the C restatement of the reference CPU path.
* ``Reference`` wraps ``oracle/_ref/libmhd_ref*.so``, the reference's own header-only solver
  compiled from its sources (present only where it was built; tests skip without it).

Both take and return the full state as a ``(9, ny_tot, nx_tot)`` float64 array in the
component order (rho, mx, my, mz, energy, bx, by, bz, psi).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "libmhd_oracle.so"
REF_DIR = ROOT / "oracle" / "_ref"

PERIODIC, REFLECTIVE, OUTFLOW = 0, 1, 2
MINMOD, MC = 0, 1
FIELDS = ("rho", "mx", "my", "mz", "energy", "bx", "by", "bz", "psi")


class Params(C.Structure):
    """Mirror of orc_params / the reference SolveParams (solve_types.hpp:8-16)."""

    _fields_ = [
        ("t_end", C.c_double),
        ("cfl", C.c_double),
        ("gamma", C.c_double),
        ("glm_alpha", C.c_double),
        ("limiter", C.c_int),
        ("ohmic", C.c_int),
        ("hall", C.c_int),
        ("ambipolar", C.c_int),
        ("eta_ohm", C.c_double),
        ("eta_hall", C.c_double),
        ("eta_ambipolar", C.c_double),
        ("max_steps", C.c_int),
    ]


def make_params(t_end=0.0, cfl=0.4, gamma=5.0 / 3.0, glm_alpha=0.1, limiter=MC,
                eta_ohm=None, eta_hall=None, eta_ambipolar=None, max_steps=1_000_000):
    p = Params()
    p.t_end, p.cfl, p.gamma, p.glm_alpha, p.limiter = t_end, cfl, gamma, glm_alpha, limiter
    p.ohmic, p.eta_ohm = (0, 0.0) if eta_ohm is None else (1, eta_ohm)
    p.hall, p.eta_hall = (0, 0.0) if eta_hall is None else (1, eta_hall)
    p.ambipolar, p.eta_ambipolar = (0, 0.0) if eta_ambipolar is None else (1, eta_ambipolar)
    p.max_steps = max_steps
    return p


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("ng", C.c_int),
        ("x_min", C.c_double),
        ("x_max", C.c_double),
        ("y_min", C.c_double),
        ("y_max", C.c_double),
    ]


def make_grid(nx, ny, x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0, ng=2):
    return Grid(nx, ny, ng, x_min, x_max, y_min, y_max)


def empty_state(g: Grid) -> np.ndarray:
    return np.zeros((9, g.ny + 2 * g.ng, g.nx + 2 * g.ng), dtype=np.float64)


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8 = C.POINTER(C.c_uint8)


def _mask_ptr(mask):
    if mask is None:
        return None, None
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    return m, m.ctypes.data_as(_u8)


class Oracle:
    """The C restatement (oracle/mhd_oracle.c)."""

    def __init__(self, path: os.PathLike = ORACLE_SO):
        if not pathlib.Path(path).exists():
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        lib = C.CDLL(str(path))
        self.lib = lib
        lib.orc_solve.restype = C.c_int
        lib.orc_solve.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, _u8, C.c_int,
                                  C.POINTER(Params), C.c_int, _dp, C.c_void_p, C.c_int,
                                  C.POINTER(C.c_double), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]
        lib.orc_solve_dt.restype = C.c_int
        lib.orc_solve_dt.argtypes = [C.POINTER(Grid), C.c_int, C.c_int, _u8, C.c_int,
                                     C.POINTER(Params), _dp, _dp, C.c_int,
                                     C.POINTER(C.c_double), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double)]
        lib.orc_init_ic.restype = C.c_int
        lib.orc_init_ic.argtypes = [C.c_char_p, C.POINTER(Grid), C.c_double, _dp]
        lib.orc_apply_bc.restype = C.c_int
        lib.orc_apply_bc.argtypes = [C.POINTER(Grid), _dp, C.c_int, C.c_int]
        lib.orc_wall_fill.restype = None
        lib.orc_wall_fill.argtypes = [C.POINTER(Grid), _dp, _u8, C.c_double]
        lib.orc_rhs.restype = None
        lib.orc_rhs.argtypes = [C.POINTER(Grid), _dp, _dp, C.POINTER(Params), C.c_double, _u8]
        lib.orc_cfl.restype = C.c_double
        lib.orc_cfl.argtypes = [C.POINTER(Grid), _dp, C.POINTER(Params), _u8,
                                C.POINTER(C.c_double)]
        lib.orc_face_flux.restype = None
        lib.orc_face_flux.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                      C.c_double, _dp]
        lib.orc_max_div_b.restype = C.c_double
        lib.orc_max_div_b.argtypes = [C.POINTER(Grid), _dp]
        lib.orc_cell_max_speed.restype = C.c_double
        lib.orc_cell_max_speed.argtypes = [_dp, C.c_double]
        lib.orc_fast_speed.restype = C.c_double
        lib.orc_fast_speed.argtypes = [_dp, C.c_double]
        lib.orc_limit.restype = C.c_double
        lib.orc_limit.argtypes = [C.c_int, C.c_double, C.c_double]
        lib.orc_nonideal_e.restype = None
        lib.orc_nonideal_e.argtypes = [_dp, _dp, C.POINTER(Params), _dp]
        lib.orc_periodic_source.restype = C.c_int
        lib.orc_periodic_source.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        lib.orc_num_threads.restype = C.c_int
        lib.orc_set_num_threads.argtypes = [C.c_int]

    def solve(self, g, U, params, bc=(PERIODIC, PERIODIC), mask=None, walls=True,
              allow_max_steps=False, dt_cap=0):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        t, steps, ch = C.c_double(), C.c_int(), C.c_double()
        log = np.zeros(max(dt_cap, 1))
        m, mp = _mask_ptr(mask)
        rc = self.lib.orc_solve(C.byref(g), bc[0], bc[1], mp, int(walls), C.byref(params),
                                int(allow_max_steps), U,
                                log.ctypes.data if dt_cap else None, dt_cap,
                                C.byref(t), C.byref(steps), C.byref(ch))
        if rc == 1:
            raise ValueError("oracle: invalid argument")
        if rc == 2:
            raise RuntimeError("oracle: reached max_steps before t_end")
        return U, dict(t=t.value, steps=steps.value, c_h=ch.value,
                       dt=log[:min(steps.value, dt_cap)] if dt_cap else None)

    def solve_dt(self, g, U, params, dts, bc=(PERIODIC, PERIODIC), mask=None, walls=True):
        """len(dts) steps with the given dt sequence (c_h from each step's CFL reduction)."""
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        d = np.ascontiguousarray(dts, dtype=np.float64)
        t, steps, ch = C.c_double(), C.c_int(), C.c_double()
        m, mp = _mask_ptr(mask)
        if self.lib.orc_solve_dt(C.byref(g), bc[0], bc[1], mp, int(walls), C.byref(params), U, d,
                                 len(d), C.byref(t), C.byref(steps), C.byref(ch)) != 0:
            raise ValueError("oracle: invalid argument")
        return U, dict(t=t.value, steps=steps.value, c_h=ch.value)

    def init_ic(self, name, g, gamma=5.0 / 3.0, bc=(PERIODIC, PERIODIC)):
        U = empty_state(g)
        if self.lib.orc_init_ic(name.encode(), C.byref(g), gamma, U) != 0:
            raise ValueError(f"unknown initial condition {name}")
        self.apply_bc(g, U, bc)
        return U

    def apply_bc(self, g, U, bc):
        if self.lib.orc_apply_bc(C.byref(g), U, bc[0], bc[1]) != 0:
            raise ValueError("oracle: bad boundary condition")
        return U

    def wall_fill(self, g, U, mask, gamma):
        m, mp = _mask_ptr(mask)
        self.lib.orc_wall_fill(C.byref(g), U, mp, gamma)
        return U

    def rhs(self, g, U, params, c_h, mask=None):
        R = np.zeros_like(U)
        m, mp = _mask_ptr(mask)
        self.lib.orc_rhs(C.byref(g), np.ascontiguousarray(U), R, C.byref(params), c_h, mp)
        return R

    def max_div_b(self, g, U):
        return self.lib.orc_max_div_b(C.byref(g), np.ascontiguousarray(U))

    def cfl(self, g, U, params, mask=None):
        ch = C.c_double()
        m, mp = _mask_ptr(mask)
        dt = self.lib.orc_cfl(C.byref(g), np.ascontiguousarray(U), C.byref(params), mp,
                              C.byref(ch))
        return dt, ch.value

    def face_flux(self, limiter, Wmm, Wm, W0, Wp, gamma, c_h):
        F = np.zeros(9)
        args = [np.ascontiguousarray(w, dtype=np.float64) for w in (Wmm, Wm, W0, Wp)]
        self.lib.orc_face_flux(limiter, *args, gamma, 1.0 / (gamma - 1.0), c_h, F)
        return F


class Reference:
    """The reference's own CPU solver compiled from its sources (oracle/ref_driver.cpp)."""

    def __init__(self, path: os.PathLike):
        lib = C.CDLL(str(path))
        self.lib = lib
        lib.ref_solve.restype = C.c_int
        lib.ref_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int, C.c_int, _u8, C.c_int, C.POINTER(Params),
                                  C.c_int, _dp, C.c_void_p, C.c_int, C.POINTER(C.c_double),
                                  C.POINTER(C.c_int), C.POINTER(C.c_double)]
        lib.ref_init_ic.restype = C.c_int
        lib.ref_init_ic.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        lib.ref_apply_bc.restype = C.c_int
        lib.ref_apply_bc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        lib.ref_wall_fill.restype = C.c_int
        lib.ref_wall_fill.argtypes = [C.c_int, C.c_int, C.c_int, _u8, C.c_double, _dp]
        lib.ref_rhs.restype = C.c_int
        lib.ref_rhs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, _dp, _dp, C.POINTER(Params), C.c_double, _u8]
        lib.ref_cfl.restype = C.c_double
        lib.ref_cfl.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, _dp, C.POINTER(Params), _u8, C.POINTER(C.c_double)]
        lib.ref_face_flux.restype = None
        lib.ref_face_flux.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                      C.c_double, _dp]
        lib.ref_max_div_b.restype = C.c_double
        lib.ref_max_div_b.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                      C.c_double, C.c_double, _dp]
        lib.ref_cell_max_speed.restype = C.c_double
        lib.ref_cell_max_speed.argtypes = [_dp, C.c_double]
        lib.ref_nonideal_e.restype = None
        lib.ref_nonideal_e.argtypes = [_dp, _dp, C.POINTER(Params), _dp]
        lib.ref_set_threads.argtypes = [C.c_int]
        lib.ref_num_threads.restype = C.c_int
        lib.ref_simd_backend.restype = C.c_char_p
        lib.ref_time_steps.restype = C.c_double
        lib.ref_time_steps.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(Params), C.c_int, _dp]

    @staticmethod
    def _g(g):
        return (g.nx, g.ny, g.ng, g.x_min, g.x_max, g.y_min, g.y_max)

    def solve(self, g, U, params, bc=(PERIODIC, PERIODIC), mask=None, walls=True,
              allow_max_steps=False, dt_cap=0):
        U = np.ascontiguousarray(U, dtype=np.float64).copy()
        t, steps, ch = C.c_double(), C.c_int(), C.c_double()
        log = np.zeros(max(dt_cap, 1))
        m, mp = _mask_ptr(mask)
        rc = self.lib.ref_solve(*self._g(g), bc[0], bc[1], mp, int(walls), C.byref(params),
                                int(allow_max_steps), U, log.ctypes.data if dt_cap else None,
                                dt_cap, C.byref(t), C.byref(steps), C.byref(ch))
        if rc == 1:
            raise ValueError("reference: invalid argument")
        if rc == 2:
            raise RuntimeError("reference: reached max_steps before t_end")
        if rc != 0:
            raise RuntimeError(f"reference: error {rc}")
        return U, dict(t=t.value, steps=steps.value, c_h=ch.value,
                       dt=log[:min(steps.value, dt_cap)] if dt_cap else None)

    def init_ic(self, name, g, gamma=5.0 / 3.0):
        U = empty_state(g)
        if self.lib.ref_init_ic(name.encode(), *self._g(g), gamma, U) != 0:
            raise ValueError(name)
        return U

    def apply_bc(self, g, U, bc):
        if self.lib.ref_apply_bc(g.nx, g.ny, g.ng, bc[0], bc[1], U) != 0:
            raise ValueError("reference: bad boundary condition")
        return U

    def wall_fill(self, g, U, mask, gamma):
        m, mp = _mask_ptr(mask)
        self.lib.ref_wall_fill(g.nx, g.ny, g.ng, mp, gamma, U)
        return U

    def rhs(self, g, U, params, c_h, mask=None):
        R = np.zeros_like(U)
        m, mp = _mask_ptr(mask)
        self.lib.ref_rhs(*self._g(g), np.ascontiguousarray(U), R, C.byref(params), c_h, mp)
        return R

    def max_div_b(self, g, U):
        return self.lib.ref_max_div_b(*self._g(g), np.ascontiguousarray(U))

    def cfl(self, g, U, params, mask=None):
        ch = C.c_double()
        m, mp = _mask_ptr(mask)
        dt = self.lib.ref_cfl(*self._g(g), np.ascontiguousarray(U), C.byref(params), mp,
                              C.byref(ch))
        return dt, ch.value

    def face_flux(self, limiter, Wmm, Wm, W0, Wp, gamma, c_h):
        F = np.zeros(9)
        args = [np.ascontiguousarray(w, dtype=np.float64) for w in (Wmm, Wm, W0, Wp)]
        self.lib.ref_face_flux(limiter, *args, gamma, 1.0 / (gamma - 1.0), c_h, F)
        return F


def exact_reference():
    """The reference built with -ffp-contract=off, or None where it was not built."""
    p = REF_DIR / "libmhd_ref.so"
    return Reference(p) if p.exists() else None


def fast_reference_path():
    """Release-flag build of the reference matching this CPU's ISA, or None."""
    try:
        flags = pathlib.Path("/proc/cpuinfo").read_text()
    except OSError:
        flags = ""
    names = ["libmhd_ref_fast_v4.so", "libmhd_ref_fast_v3.so"] if " avx512f" in flags \
        else ["libmhd_ref_fast_v3.so"] if " avx2" in flags else []
    for n in names:
        if (REF_DIR / n).exists():
            return REF_DIR / n
    return None


def read_golden(path):
    """Reader for the reference's text golden format (tests/golden_io.hpp:12-81)."""
    meta, fields = {}, {}
    with open(path) as f:
        tok = f.read().split()
    assert tok[0] == "MHD_GOLDEN" and tok[1] == "1", path
    k = 2
    while k < len(tok):
        key = tok[k]
        if key == "field":
            name, n = tok[k + 1], int(tok[k + 2])
            fields[name] = np.array([float(v) for v in tok[k + 3:k + 3 + n]])
            k += 3 + n
        else:
            meta[key] = tok[k + 1]
            k += 2
    return meta, fields


def read_ref_scalars(path):
    """Reader for the reference's MHD_REF scalar files."""
    tok = open(path).read().split()
    assert tok[0] == "MHD_REF" and tok[1] == "1"
    return {tok[i]: float(tok[i + 1]) for i in range(2, len(tok), 2)}


def interior(U, ng=2):
    return U[..., ng:-ng, ng:-ng]
