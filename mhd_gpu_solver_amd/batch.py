"""Batched ensemble engine (include/mhd_hip.h, mhd_hip_batch_*).

Many independent solves of different grid sizes stay resident on the GPU and advance together,
one persistent kernel per RK stage, each with its own dt. Each instance follows the ::solve()
contract of the reference (include/mhd/time_integrator.hpp:62-106) and its result does not
depend on which other instances share the batch.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _lib
from .solver import NonIdealConfig, SlopeLimiter, SolveResult

CapacityFull = _lib.CapacityFull


def instance_bytes(precision: str, nx: int, ny: int) -> int:
    """HBM bytes one nx x ny instance occupies (U and U*)."""
    return int(_lib.lib().mhd_hip_batch_instance_bytes(_prec(precision), int(nx), int(ny)))


def _prec(precision):
    if precision not in _lib.PRECISIONS:
        raise ValueError("precision must be 'fp64', 'fp32' or 'fp64_fast'")
    return _lib.PRECISIONS[precision]


class BatchEngine:
    def __init__(self, precision="fp32", limiter=SlopeLimiter.MC, nonideal=None,
                 hbm_budget=8 << 30, max_instances=1024):
        self._l = _lib.lib()
        p = _lib.Params()
        p.limiter = int(limiter)
        ni = nonideal or NonIdealConfig()
        p.ohmic, p.hall, p.ambipolar = int(ni.ohmic), int(ni.hall), int(ni.ambipolar)
        p.eta_ohm, p.eta_hall, p.eta_ambipolar = ni.eta_ohm, ni.eta_hall, ni.eta_ambipolar
        h = C.c_void_p()
        _lib.check(self._l.mhd_hip_batch_create(_prec(precision), C.byref(p), int(hbm_budget),
                                                int(max_instances), C.byref(h)))
        self.handle = h
        self.precision = precision
        self.limiter = SlopeLimiter(limiter)
        self.nonideal = dataclasses.replace(ni)
        self._shape = {}
        self._ids = (C.c_int * int(max_instances))()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self._l.mhd_hip_batch_destroy(h)
            self.handle = None

    def add(self, state, t_end, cfl=0.4, gamma=5.0 / 3.0, glm_alpha=0.1, bc=(0, 0),
            max_steps=1_000_000, x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0, limiter=None,
            nonideal=None, mask=None, walls=True) -> int:
        """Admit one solve; `state` is (9, ny+4, nx+4) float64 (ghosts included).
        `limiter` and `nonideal` default to the engine's; `mask` is an optional (ny, nx) uint8
        FluidMask (1 = fluid) and `walls` its reflective wall fill, as MHDSolver.set_mask /
        set_walls_reflective. Raises CapacityFull when the HBM budget or the slots are
        exhausted."""
        U = np.ascontiguousarray(state, dtype=np.float64)
        ny, nx = U.shape[1] - 4, U.shape[2] - 4
        d = _lib.Instance(nx, ny, x_min, x_max, y_min, y_max, int(bc[0]), int(bc[1]),
                          float(t_end), float(cfl), float(gamma), float(glm_alpha),
                          int(max_steps))
        ptrs, _ = _lib.planes(U)
        i = C.c_int()
        if limiter is None and nonideal is None and mask is None:
            _lib.check(self._l.mhd_hip_batch_add(self.handle, C.byref(d), ptrs, C.byref(i)))
        else:
            p = _lib.Params()
            p.limiter = int(self.limiter if limiter is None else limiter)
            ni = self.nonideal if nonideal is None else nonideal
            p.ohmic, p.hall, p.ambipolar = int(ni.ohmic), int(ni.hall), int(ni.ambipolar)
            p.eta_ohm, p.eta_hall, p.eta_ambipolar = ni.eta_ohm, ni.eta_hall, ni.eta_ambipolar
            if mask is None:
                mptr, mlen = None, 0
            else:
                m = np.ascontiguousarray(mask, dtype=np.uint8)
                if m.shape != (ny, nx):
                    raise ValueError("mask shape must be (ny, nx)")
                mptr, mlen = m.ctypes.data_as(C.POINTER(C.c_uint8)), m.size
            _lib.check(self._l.mhd_hip_batch_add_config(self.handle, C.byref(d), C.byref(p), mptr,
                                                        mlen, int(bool(walls)), ptrs,
                                                        C.byref(i)))
        self._shape[i.value] = (ny + 4, nx + 4)
        return i.value

    def add_steps(self, state, steps, **kw) -> int:
        """Admit a fixed-step-count solve (t_end = +inf, max_steps = steps)."""
        return self.add(state, t_end=math.inf, max_steps=steps, **kw)

    def remove(self, iid: int):
        _lib.check(self._l.mhd_hip_batch_remove(self.handle, int(iid)))
        self._shape.pop(iid, None)

    def prepare(self):
        """Plan the next replay now (tile tables, timed row choice, step graph), so that the
        next advance starts with its launch."""
        _lib.check(self._l.mhd_hip_batch_prepare(self.handle))

    def advance(self, max_replays: int = -1) -> int:
        """Step until no instance is active (or max_replays graph replays); returns active."""
        a = C.c_int()
        _lib.check(self._l.mhd_hip_batch_advance(self.handle, int(max_replays), C.byref(a)))
        return a.value

    def result(self, iid: int):
        r = _lib.Result()
        f = C.c_int()
        _lib.check(self._l.mhd_hip_batch_result(self.handle, int(iid), C.byref(r), C.byref(f)))
        return SolveResult(r.t, r.steps, r.c_h), bool(f.value)

    def download(self, iid: int):
        shape = self._shape[iid]
        U = np.zeros((9,) + shape)
        ptrs, n = _lib.planes(U, writable=True)
        _lib.check(self._l.mhd_hip_batch_download(self.handle, int(iid), ptrs, n))
        return U

    def totals(self, iid: int):
        """Compensated device sums of the nine conserved components of one instance."""
        out = (C.c_double * 9)()
        _lib.check(self._l.mhd_hip_batch_totals(self.handle, int(iid), out))
        return np.array(out[:])

    def rel_l1(self, iid: int, ref: "BatchEngine", ref_iid: int):
        """Relative L1 distance per conserved component of instance `iid` from instance
        `ref_iid` of engine `ref` (this engine or another, e.g. fp64 against FP32), over the
        interior, computed on the device (DeviceSolver.rel_l1)."""
        out = (C.c_double * 9)()
        _lib.check(self._l.mhd_hip_batch_rel_l1(self.handle, int(iid), ref.handle, int(ref_iid),
                                                out))
        return np.array(out[:])

    def max_div_b(self, iid: int) -> float:
        """max |div B| of one instance on the device (after its final ghost fill)."""
        v = C.c_double()
        _lib.check(self._l.mhd_hip_batch_max_div_b(self.handle, int(iid), C.byref(v)))
        return v.value

    def set_seg_rows(self, rows: int):
        """Work partition of the batch kernel: 0 (default) rows per tile from the load model
        at every change of the stepping set, -1 the fixed lengths the load model ranks best timed
        on dry steps (no instance advances) before the first replay and the fastest kept,
        > 0 fixed rows per tile, < -1 the capacity fill with pieces of at least -rows rows.
        Results do not depend on it."""
        _lib.check(self._l.mhd_hip_batch_set_seg_rows(self.handle, int(rows)))

    def set_paused(self, iid: int, paused: bool):
        """Keep an instance resident but out of the stepping set (from the next replay)."""
        _lib.check(self._l.mhd_hip_batch_set_paused(self.handle, int(iid), int(paused)))

    def progress(self, iid: int):
        """(t, steps, dt estimate) as of the last replay."""
        t, n, dt = C.c_double(), C.c_int(), C.c_double()
        _lib.check(self._l.mhd_hip_batch_progress(self.handle, int(iid), C.byref(t), C.byref(n),
                                                  C.byref(dt)))
        return t.value, n.value, dt.value

    SCHEDULES = {"all": 0, "srpt": 1}

    def set_schedule(self, policy: str, work_cap_cells: int = 1):
        """"all": every unfinished solve steps each replay. "srpt": before each replay the
        engine steps the least-remaining-work solves until `work_cap_cells` cells step and
        pauses the rest (it then owns the paused flags)."""
        if policy not in self.SCHEDULES:
            raise ValueError(f"unknown scheduling policy {policy!r}")
        _lib.check(self._l.mhd_hip_batch_set_schedule(self.handle, self.SCHEDULES[policy],
                                                      int(work_cap_cells)))

    def poll_finished(self):
        """Ids of the finished solves still resident, as of the last replay."""
        n = C.c_int()
        _lib.check(self._l.mhd_hip_batch_poll_finished(self.handle, self._ids, len(self._ids),
                                                       C.byref(n)))
        return list(self._ids[:n.value])

    def set_kernel_timing(self, enable: bool):
        _lib.check(self._l.mhd_hip_batch_set_kernel_timing(self.handle, int(enable)))

    def kernel_times(self):
        """[(total_ms, launches) of the stage-1 kernel, same for stage 2]"""
        ms = (C.c_double * 2)()
        n = (C.c_longlong * 2)()
        _lib.check(self._l.mhd_hip_batch_kernel_times(self.handle, ms, n))
        return [(ms[0], n[0]), (ms[1], n[1])]

    def stats(self) -> dict:
        st = _lib.BatchStats()
        _lib.check(self._l.mhd_hip_batch_info(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in _lib.BatchStats._fields_}
