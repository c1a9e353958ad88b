"""Python mirror of the reference's ``mhd_solver`` module (python/mhd_pybind.cpp:166-309) with
the ``hip`` backend behind it.

``MHDSolver`` keeps the reference's owning-API semantics (include/mhd/solver_api.hpp:22-452):
the host state (a ``(9, ny_tot, nx_tot)`` float64 array, ghosts included) is authoritative
between calls, and ``run`` does configure -> upload -> solve -> download on the device solver
(solver_api.hpp:294-321), accumulating ``time`` and ``steps``. Every time step runs in the HIP
kernels; there is no CPU solver in this package.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import enum

import numpy as np

from . import _lib

FIELDS = ("rho", "mx", "my", "mz", "energy", "bx", "by", "bz", "psi")


class SlopeLimiter(enum.IntEnum):
    """include/mhd/mhd_types.hpp:44"""
    Minmod = 0
    MC = 1


class BoundaryConditionType(enum.IntEnum):
    """include/boundary_conditions/boundary_conditions.hpp:8"""
    Periodic = 0
    Reflective = 1
    Outflow = 2


@dataclasses.dataclass
class NonIdealConfig:
    """include/physics_config/mhd_config.hpp:13-111"""
    ohmic: bool = False
    hall: bool = False
    ambipolar: bool = False
    eta_ohm: float = 0.0
    eta_hall: float = 0.0
    eta_ambipolar: float = 0.0

    @staticmethod
    def ideal():
        return NonIdealConfig()

    @staticmethod
    def with_ohmic(eta):
        return NonIdealConfig(ohmic=True, eta_ohm=eta)

    @staticmethod
    def with_hall(eta):
        return NonIdealConfig(hall=True, eta_hall=eta)

    @staticmethod
    def with_ambipolar(eta):
        return NonIdealConfig(ambipolar=True, eta_ambipolar=eta)

    def enable_ohmic(self, eta):
        self.ohmic, self.eta_ohm = True, eta
        return self

    def enable_hall(self, eta):
        self.hall, self.eta_hall = True, eta
        return self

    def enable_ambipolar(self, eta):
        self.ambipolar, self.eta_ambipolar = True, eta
        return self

    def disable_ohmic(self):
        self.ohmic = False
        return self

    def disable_hall(self):
        self.hall = False
        return self

    def disable_ambipolar(self):
        self.ambipolar = False
        return self

    def any(self):
        return ((self.ohmic and self.eta_ohm != 0.0) or (self.hall and self.eta_hall != 0.0)
                or (self.ambipolar and self.eta_ambipolar != 0.0))

    def max_diffusivity(self):
        eta = 0.0
        for on, v in ((self.ohmic, self.eta_ohm), (self.hall, self.eta_hall),
                      (self.ambipolar, self.eta_ambipolar)):
            if on:
                eta = max(eta, abs(v))
        return eta


@dataclasses.dataclass
class SolveParams:
    """include/mhd/solve_types.hpp:8-16"""
    t_end: float = 0.0
    cfl: float = 0.4
    gamma: float = 5.0 / 3.0
    glm_alpha: float = 0.1
    limiter: SlopeLimiter = SlopeLimiter.MC
    nonideal: NonIdealConfig = dataclasses.field(default_factory=NonIdealConfig)
    max_steps: int = 1_000_000

    def to_c(self) -> _lib.Params:
        p = _lib.Params()
        p.t_end, p.cfl, p.gamma, p.glm_alpha = self.t_end, self.cfl, self.gamma, self.glm_alpha
        p.limiter = int(self.limiter)
        ni = self.nonideal
        p.ohmic, p.hall, p.ambipolar = int(ni.ohmic), int(ni.hall), int(ni.ambipolar)
        p.eta_ohm, p.eta_hall, p.eta_ambipolar = ni.eta_ohm, ni.eta_hall, ni.eta_ambipolar
        p.max_steps = int(self.max_steps)
        return p


@dataclasses.dataclass(frozen=True)
class SolveResult:
    """include/mhd/solve_types.hpp:18-22"""
    t: float
    steps: int
    c_h: float


def hip_available() -> bool:
    return bool(_lib.lib().mhd_hip_available())


def stream_bandwidth(precision: str, cells: int, stage: int, reps: int = 20) -> float:
    """GB/s of a streaming kernel with one RK stage's memory traffic and none of its arithmetic
    (stage 1: 9 planes read, 9 written; stage 2: 18 read, 9 written in place; `cells` elements
    per plane), measured on the current device (include/mhd_hip.h, mhd_hip_stream_bandwidth).
    What the memory system delivers for the stage's mix, beside the published HBM peak."""
    if precision not in _lib.PRECISIONS:
        raise ValueError("precision must be 'fp64', 'fp32' or 'fp64_fast'")
    v = C.c_double()
    _lib.check(_lib.lib().mhd_hip_stream_bandwidth(_lib.PRECISIONS[precision], int(cells),
                                                   int(stage), int(reps), C.byref(v)))
    return v.value


def cuda_available() -> bool:
    return False


def metal_available() -> bool:
    return False


def available_backends():
    """Backends usable here: this package provides ``hip`` only (python/mhd_pybind.cpp:185-195)."""
    return ["hip"] if hip_available() else []


class DeviceSolver:
    """Owner of one ``mhd_hip_solver`` handle (the DeviceSolver of device_solver.hpp:57-75)."""

    def __init__(self, grid: _lib.Grid, precision: int):
        self._l = _lib.lib()
        h = C.c_void_p()
        _lib.check(self._l.mhd_hip_create(C.byref(grid), precision, C.byref(h)), "create")
        self.handle = h
        self.grid = grid

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self._l.mhd_hip_destroy(h)
            self.handle = None

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        _lib.check(self._l.mhd_hip_device_name(self.handle, buf, 256))
        return buf.value.decode()

    def precision(self) -> str:
        code = self._l.mhd_hip_precision(self.handle)
        return next(k for k, v in _lib.PRECISIONS.items() if v == code)

    def configure(self, bc_x, bc_y, mask, reflective_walls):
        if mask is None:
            _lib.check(self._l.mhd_hip_configure(self.handle, int(bc_x), int(bc_y), None, 0,
                                                 int(reflective_walls)))
        else:
            m = np.ascontiguousarray(mask, dtype=np.uint8)
            _lib.check(self._l.mhd_hip_configure(self.handle, int(bc_x), int(bc_y),
                                                 m.ctypes.data_as(C.POINTER(C.c_uint8)), m.size,
                                                 int(reflective_walls)))

    def upload(self, U):
        ptrs, n = _lib.planes(U)
        _lib.check(self._l.mhd_hip_upload(self.handle, ptrs, n))

    def download(self, U):
        ptrs, n = _lib.planes(U, writable=True)
        _lib.check(self._l.mhd_hip_download(self.handle, ptrs, n))

    def solve(self, params: SolveParams) -> SolveResult:
        r = _lib.Result()
        _lib.check(self._l.mhd_hip_solve(self.handle, C.byref(params.to_c()), C.byref(r)))
        return SolveResult(r.t, r.steps, r.c_h)

    def advance_steps(self, params: SolveParams, steps: int) -> SolveResult:
        r = _lib.Result()
        _lib.check(self._l.mhd_hip_advance_steps(self.handle, C.byref(params.to_c()), int(steps),
                                                 C.byref(r)))
        return SolveResult(r.t, r.steps, r.c_h)

    def set_dt_log(self, capacity: int):
        _lib.check(self._l.mhd_hip_set_dt_log(self.handle, int(capacity)))

    def dt_log(self, capacity: int):
        out = np.zeros(max(1, capacity))
        cnt = C.c_int()
        _lib.check(self._l.mhd_hip_dt_log(self.handle, out.ctypes.data_as(C.POINTER(C.c_double)),
                                          int(capacity), C.byref(cnt)))
        return out[:cnt.value]

    def fill_boundaries(self, gamma: float, ghosts=True, walls=True):
        what = (_lib.FILL_GHOSTS if ghosts else 0) | (_lib.FILL_WALLS if walls else 0)
        _lib.check(self._l.mhd_hip_fill_boundaries(self.handle, gamma, what))

    def totals(self):
        """Compensated device sums of the nine conserved components over the interior."""
        out = (C.c_double * 9)()
        _lib.check(self._l.mhd_hip_totals(self.handle, out))
        return np.array(out[:])

    def max_div_b(self) -> float:
        """max |div B| of the device state (central differences, ghosts included), on the
        device; equals MHDSolver.max_div_b of the same state."""
        v = C.c_double()
        _lib.check(self._l.mhd_hip_max_div_b(self.handle, C.byref(v)))
        return v.value

    def rel_l1(self, ref: "DeviceSolver"):
        """Relative L1 distance per conserved component of this device state from `ref`'s
        (same grid, either precision), over the interior, computed on the device:
        sum |a - ref| / sum |ref| (the absolute sum where sum |ref| is 0)."""
        out = (C.c_double * 9)()
        _lib.check(self._l.mhd_hip_rel_l1(self.handle, ref.handle, out))
        return np.array(out[:])

    def set_seg_rows(self, rows: int):
        """Rows marched per wavefront (0 = autotuned); results do not depend on it."""
        _lib.check(self._l.mhd_hip_set_seg_rows(self.handle, int(rows)))

    def seg_rows(self) -> int:
        v = C.c_int()
        _lib.check(self._l.mhd_hip_get_seg_rows(self.handle, C.byref(v)))
        return v.value

    def set_kernel_timing(self, enable: bool):
        _lib.check(self._l.mhd_hip_set_kernel_timing(self.handle, int(enable)))

    def kernel_times(self):
        """[(total_ms, launches) for the stage-1 kernel, same for stage 2]"""
        ms = (C.c_double * 2)()
        n = (C.c_longlong * 2)()
        _lib.check(self._l.mhd_hip_kernel_times(self.handle, ms, n))
        return [(ms[0], n[0]), (ms[1], n[1])]

    def last_solve_seconds(self) -> float:
        s = C.c_double()
        _lib.check(self._l.mhd_hip_last_solve_seconds(self.handle, C.byref(s)))
        return s.value


def init_ic(name: str, nx: int, ny: int, x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0, ng=2,
            gamma=5.0 / 3.0, b0=1.0, seed=0, out=None):
    """Host-side initial condition into the interior of a (9, ny_tot, nx_tot) state (CPU only)."""
    g = _lib.Grid(nx, ny, ng, x_min, x_max, y_min, y_max)
    U = np.zeros((9, ny + 2 * ng, nx + 2 * ng)) if out is None else out
    ptrs, _ = _lib.planes(U, writable=True)
    _lib.check(_lib.lib().mhd_hip_init_ic(name.encode(), C.byref(g), gamma, b0, seed, ptrs))
    return U


def wall_plan(mask, nx, ny, ng=2):
    """Solid-cell fill plan (dst, src or -1, axis) as an (n, 3) int64 array (CPU only)."""
    g = _lib.Grid(nx, ny, ng, 0.0, 1.0, 0.0, 1.0)
    m = np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    cnt = C.c_size_t()
    _lib.check(_lib.lib().mhd_hip_wall_plan(C.byref(g), m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            m.size, None, 0, C.byref(cnt)))
    plan = np.zeros(3 * max(1, cnt.value), dtype=np.int64)
    _lib.check(_lib.lib().mhd_hip_wall_plan(C.byref(g), m.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            m.size, plan.ctypes.data_as(C.POINTER(C.c_int64)),
                                            plan.size, C.byref(cnt)))
    return plan[:3 * cnt.value].reshape(-1, 3)


class MHDSolver:
    """Owning solver API (include/mhd/solver_api.hpp:22-452) on the hip backend."""

    def __init__(self, nx, ny, x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0, ng=2):
        if ng < 2:
            raise ValueError("MHDSolver requires ng >= 2")
        self._grid = _lib.Grid(int(nx), int(ny), int(ng), float(x_min), float(x_max),
                               float(y_min), float(y_max))
        self._U = np.zeros((9, ny + 2 * ng, nx + 2 * ng))
        self._mask = np.ones((ny, nx), dtype=np.uint8)
        self._bc = (BoundaryConditionType.Periodic, BoundaryConditionType.Periodic)
        self.params = SolveParams()
        self._t = 0.0
        self._steps = 0
        self._c_h = 0.0
        self._walls = True
        self._precision = "fp64"
        self._device = None

    # ---- grid -------------------------------------------------------------------------------
    nx = property(lambda self: self._grid.nx)
    ny = property(lambda self: self._grid.ny)
    ng = property(lambda self: self._grid.ng)
    size_x = property(lambda self: self._grid.nx + 2 * self._grid.ng)
    size_y = property(lambda self: self._grid.ny + 2 * self._grid.ng)
    x_min = property(lambda self: self._grid.x_min)
    x_max = property(lambda self: self._grid.x_max)
    y_min = property(lambda self: self._grid.y_min)
    y_max = property(lambda self: self._grid.y_max)
    dx = property(lambda self: (self._grid.x_max - self._grid.x_min) / self._grid.nx)
    dy = property(lambda self: (self._grid.y_max - self._grid.y_min) / self._grid.ny)

    # ---- configuration (solver_api.hpp:80-145) --------------------------------------------------
    def set_periodic_bc(self):
        self._bc = (BoundaryConditionType.Periodic, BoundaryConditionType.Periodic)

    def set_outflow_bc(self):
        self._bc = (BoundaryConditionType.Outflow, BoundaryConditionType.Outflow)

    def set_bc(self, type_x, type_y):
        self._bc = (BoundaryConditionType(type_x), BoundaryConditionType(type_y))

    def set_gamma(self, gamma):
        self.params.gamma = float(gamma)

    def set_cfl(self, cfl):
        self.params.cfl = float(cfl)

    def set_glm_alpha(self, alpha):
        self.params.glm_alpha = float(alpha)

    def set_limiter_minmod(self):
        self.params.limiter = SlopeLimiter.Minmod

    def set_limiter_mc(self):
        self.params.limiter = SlopeLimiter.MC

    def set_ideal(self):
        self.params.nonideal = NonIdealConfig.ideal()

    def enable_ohmic(self, eta):
        self.params.nonideal.enable_ohmic(eta)

    def enable_hall(self, eta):
        self.params.nonideal.enable_hall(eta)

    def enable_ambipolar(self, eta):
        self.params.nonideal.enable_ambipolar(eta)

    def disable_ohmic(self):
        self.params.nonideal.disable_ohmic()

    def disable_hall(self):
        self.params.nonideal.disable_hall()

    def disable_ambipolar(self):
        self.params.nonideal.disable_ambipolar()

    # ---- mask (solver_api.hpp:148-163, mhd_pybind.cpp:50-69) -------------------------------------
    def set_mask(self, mask):
        m = np.asarray(mask)
        if m.ndim != 2:
            raise ValueError("set_mask expects a 2D array (ny, nx)")
        if m.shape != (self.ny, self.nx):
            raise ValueError("set_mask shape must be (ny, nx)")
        self._mask = np.ascontiguousarray(m, dtype=np.uint8)
        if self._walls:
            self._fill_boundaries(ghosts=False)

    def get_mask(self):
        return self._mask.copy()

    def clear_mask(self):
        self._mask = np.ones((self.ny, self.nx), dtype=np.uint8)

    def set_walls_reflective(self, enabled):
        self._walls = bool(enabled)

    @property
    def walls_reflective(self):
        return self._walls

    def _active_mask(self):
        return None if bool(self._mask.all()) else self._mask

    # ---- state I/O (solver_api.hpp:211-274, mhd_pybind.cpp:72-162) ------------------------------
    def _index(self, name):
        if name not in FIELDS:
            raise ValueError(f"Unknown field: {name}")
        return FIELDS.index(name)

    def set_field(self, name, array):
        c = self._index(name)
        a = np.asarray(array, dtype=np.float64)
        if a.ndim != 2:
            raise ValueError("set_field expects a 2D array")
        ng = self.ng
        if a.shape == (self.size_y, self.size_x):
            self._U[c] = a
        elif a.shape == (self.ny, self.nx):
            self._U[c, ng:ng + self.ny, ng:ng + self.nx] = a
        else:
            raise ValueError("set_field shape must be full (size_y, size_x) or interior (ny, nx)")

    def field(self, name):
        return self._U[self._index(name)].copy()

    def copy_field(self, name):
        return self._U[self._index(name)].ravel().copy()

    @property
    def state(self):
        """The host state, shape (9, size_y, size_x), component order rho..psi (a live view)."""
        return self._U

    def set_primitives(self, rho, vx, vy, vz, pressure, bx, by, bz, psi=None):
        arrs = [np.asarray(a, dtype=np.float64) for a in (rho, vx, vy, vz, pressure, bx, by, bz)]
        for a in arrs:
            if a.ndim != 2:
                raise ValueError("primitive arrays must be 2D")
            if a.shape != (self.ny, self.nx):
                raise ValueError("primitive arrays must have shape (ny, nx)")
        rho, vx, vy, vz, p, bx, by, bz = arrs
        psi = np.zeros_like(rho) if psi is None else np.asarray(psi, dtype=np.float64)
        g = self.params.gamma
        # to_conservative: include/mhd/mhd_physics.hpp:41-59 (same operation order)
        e_kin = 0.5 * rho * (vx * vx + vy * vy + vz * vz)
        e_mag = 0.5 * (bx * bx + by * by + bz * bz)
        cons = (rho, rho * vx, rho * vy, rho * vz, p / (g - 1.0) + e_kin + e_mag, bx, by, bz, psi)
        ng = self.ng
        for c, v in enumerate(cons):
            self._U[c, ng:ng + self.ny, ng:ng + self.nx] = v
        self._fill_boundaries(ghosts=True)

    def initialize(self, name, b0=1.0, seed=0):
        """Named initial condition (solver_api.hpp:276-292). ``b0``/``seed`` parametrise this
        backend's workload ICs (orszag_tang_b0, rotor_b0, kh_tanh, alfven_oblique)."""
        if name in ("sod", "Sod"):
            self.params.gamma = 1.4
        elif name in ("brio_wu", "BrioWu"):
            self.params.gamma = 2.0
        g = self._grid
        init_ic(name, g.nx, g.ny, g.x_min, g.x_max, g.y_min, g.y_max, g.ng, self.params.gamma,
                b0, seed, out=self._U)
        self._fill_boundaries(ghosts=True)
        self._t = 0.0
        self._steps = 0

    def _fill_boundaries(self, ghosts: bool):
        """Ghost fill (+ wall fill) of the host state, done by the device kernels."""
        if not ghosts and self._active_mask() is None:
            return
        dev = self._dev()
        dev.configure(self._bc[0], self._bc[1], self._active_mask(), self._walls)
        dev.upload(self._U)
        dev.fill_boundaries(self.params.gamma, ghosts=ghosts, walls=True)
        dev.download(self._U)

    # ---- backend (solver_api.hpp:169-198) --------------------------------------------------------
    def set_backend(self, backend="hip", precision="fp64"):
        if backend in ("hip", "HIP"):
            if not hip_available():
                raise RuntimeError("Backend not available: hip")
        elif backend in ("cpu", "CPU", "cuda", "CUDA", "metal", "Metal"):
            raise RuntimeError(f"Backend not available: {backend.lower()} "
                               "(this package provides the hip backend)")
        else:
            raise ValueError(f"Unknown backend: {backend} (expected hip)")
        # solver_api.hpp:180-184: "fp32" selects FP32 and any other string fp64; "fp64_fast"
        # (this backend's contracted fp64 build) is the one name added
        self._precision = precision if precision in ("fp32", "fp64_fast") else "fp64"
        self._device = None

    @property
    def backend(self):
        return "hip"

    @property
    def precision(self):
        return self._precision

    def backend_device(self):
        return self._dev().device_name()

    def _dev(self) -> DeviceSolver:
        if self._device is None:
            self._device = DeviceSolver(self._grid, _lib.PRECISIONS[self._precision])
        return self._device

    # ---- time integration (solver_api.hpp:294-329) -----------------------------------------------
    def run(self, t_end) -> SolveResult:
        if t_end < 0.0:
            raise ValueError("run: t_end must be non-negative")
        local = dataclasses.replace(self.params, t_end=float(t_end),
                                    nonideal=dataclasses.replace(self.params.nonideal))
        dev = self._dev()
        dev.configure(self._bc[0], self._bc[1], self._active_mask(), self._walls)
        dev.upload(self._U)
        r = dev.solve(local)
        dev.download(self._U)
        self._t += r.t
        self._steps += r.steps
        self._c_h = r.c_h
        return SolveResult(self._t, self._steps, self._c_h)

    def advance_to(self, t_abs) -> SolveResult:
        if t_abs < self._t:
            raise ValueError("advance_to: target time is earlier than current time")
        return self.run(t_abs - self._t)

    def advance_steps(self, steps) -> SolveResult:
        """Exactly ``steps`` steps (no t_end); for benchmarks and dt-sequence checks."""
        dev = self._dev()
        dev.configure(self._bc[0], self._bc[1], self._active_mask(), self._walls)
        dev.upload(self._U)
        r = dev.advance_steps(self.params, steps)
        dev.download(self._U)
        self._t += r.t
        self._steps += r.steps
        self._c_h = r.c_h
        return SolveResult(self._t, self._steps, self._c_h)

    time = property(lambda self: self._t)
    steps = property(lambda self: self._steps)
    c_h = property(lambda self: self._c_h)

    def conserved_totals(self):
        """Integrals of the nine conserved components over the domain (sum * dx * dy), summed
        on the device in double-double arithmetic from the current host state."""
        dev = self._dev()
        dev.configure(self._bc[0], self._bc[1], self._active_mask(), self._walls)
        dev.upload(self._U)
        return dev.totals() * (self.dx * self.dy)

    def max_div_b(self):
        """include/initial_condition/inititial_condition.hpp:273-287"""
        ng, nx, ny = self.ng, self.nx, self.ny
        bx, by = self._U[5], self._U[6]
        dbx = (bx[ng:ng + ny, ng + 1:ng + nx + 1] - bx[ng:ng + ny, ng - 1:ng + nx - 1]) * (0.5 / self.dx)
        dby = (by[ng + 1:ng + ny + 1, ng:ng + nx] - by[ng - 1:ng + ny - 1, ng:ng + nx]) * (0.5 / self.dy)
        return float(np.max(np.abs(dbx + dby))) if nx * ny else 0.0
