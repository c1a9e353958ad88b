"""Continuous-batching request loop over the batched engine.

Solve requests (initial condition, grid, gamma, boundaries, t_end) arrive over time. Arrived
requests wait in a FIFO queue and are admitted into the running batch as soon as its HBM budget
and instance slots allow; the engine advances everything admitted by one graph replay at a
time; finished solves are downloaded, released and their latency (completion - arrival)
recorded. New requests join between replays while the others keep running, so their work
overlaps on the device. No counterpart exists in the reference, which solves one grid per
call (include/mhd/solver_api.hpp:294-321).

Host preparation: the initial state of an arrived request is built on a few host threads while
the device runs the current replay (up to ``prepare_ahead`` requests ahead of admission), so
the device does not wait on it. Admission keeps FIFO order: while solves are running, a queue
head whose state is not ready yet waits for the next replay.

Scheduling: with ``policy="all"`` every admitted solve steps in every replay (processor
sharing). With ``policy="srpt"`` admitted solves stay resident in HBM but each replay steps
only those with the least estimated remaining work (cells x remaining steps, the steps taken
from the device's own dt so far) until ``work_cap_cells`` cells are stepping; the rest are
paused. Pausing never changes a solve's result, only when it is computed.
"""
from __future__ import annotations

import collections
import dataclasses
import itertools
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .batch import BatchEngine, CapacityFull, instance_bytes
from .solver import SlopeLimiter


def _scheme(req) -> dict:
    """A request's own limiter, non-ideal terms and mask, where it names them."""
    out = {}
    for key in ("limiter", "nonideal", "mask"):
        if getattr(req, key, None) is not None:
            out[key] = getattr(req, key)
    if "mask" in out:
        out["walls"] = getattr(req, "walls", True)
    return out


@dataclasses.dataclass
class Completed:
    index: int
    n: int
    ic: str
    arrival: float
    admitted: float
    finished: float
    steps: int
    t: float
    state: np.ndarray | None = None
    error: str | None = None  # the reference's solve() error for this request, if any

    @property
    def latency(self):
        return self.finished - self.arrival

    @property
    def cell_updates(self):
        return self.n * self.n * self.steps


@dataclasses.dataclass
class ServeReport:
    completed: list
    makespan: float
    max_concurrent: int
    replays: int
    advance_s: float = 0.0  # wall time inside the engine's replays (the device's busy time)

    def latencies_ms(self):
        return np.array([c.latency for c in self.completed if c.error is None]) * 1e3

    def summary(self) -> dict:
        lat = self.latencies_ms()
        cu = sum(c.cell_updates for c in self.completed)
        return {
            "requests": len(self.completed),
            "failed": sum(1 for c in self.completed if c.error is not None),
            "p50_ms": round(float(np.percentile(lat, 50)), 2) if len(lat) else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 2) if len(lat) else None,
            "mean_ms": round(float(lat.mean()), 2) if len(lat) else None,
            "makespan_s": round(self.makespan, 3),
            "mcups": round(cu / self.makespan / 1e6, 2) if self.makespan > 0 else None,
            "cell_updates": cu,
            "max_concurrent": self.max_concurrent,
            "replays": self.replays,
            "advance_s": round(self.advance_s, 3),
        }


class Server:
    # Cells stepping per replay under srpt. With tiles handed out from per-XCD queues, the
    # request stream of 128 solves at 200/s (FP32, MI355X) keeps its throughput down to about
    # 5M cells and loses 5 % at 4M, while the median latency falls with the cap: 8M 0.67-0.72 s at 19.0-19.2 G cell-updates/s, p99
    # 6.39-6.42 s; 6M 0.58 s, 19.4 G, p99 6.19 s; 5M 0.48 s, 19.4 G, p99 6.23 s; 4M 0.48 s,
    # 18.3 G, p99 6.38 s.
    DEFAULT_WORK_CAP = 5 << 20

    def __init__(self, precision="fp32", limiter=SlopeLimiter.MC, hbm_budget=32 << 30,
                 max_instances=1024, cfl=0.4, glm_alpha=0.1, policy="srpt",
                 work_cap_cells=DEFAULT_WORK_CAP):
        if policy not in ("all", "srpt"):
            raise ValueError(f"unknown scheduling policy {policy!r}")
        self.engine = BatchEngine(precision, limiter=limiter, hbm_budget=hbm_budget,
                                  max_instances=max_instances)
        self.precision = precision
        self.cfl = cfl
        self.glm_alpha = glm_alpha
        self.hbm_budget = hbm_budget
        self.policy = policy
        self.work_cap_cells = int(work_cap_cells)
        self.engine.set_schedule(policy, self.work_cap_cells)

    def run(self, requests, realtime=True, keep_states=False, clock=time.perf_counter,
            prepare_workers=4, prepare_ahead=16):
        """Serve `requests` (workloads.Spec with .arrival seconds, .t_end) to completion.
        realtime: arrivals follow the wall clock from the start of the call; otherwise every
        request is available at once. Initial states are built by `prepare_workers` host
        threads, at most `prepare_ahead` arrived requests ahead of admission."""
        with ThreadPoolExecutor(max_workers=max(1, prepare_workers)) as prep:
            return self._run(requests, realtime, keep_states, clock, prep,
                             max(1, prepare_ahead))

    def _run(self, requests, realtime, keep_states, clock, prep, prepare_ahead):
        order = sorted(range(len(requests)), key=lambda k: requests[k].arrival)
        pending = collections.deque(order)
        queue = collections.deque()
        states = {}  # request index -> future of its initial state
        active = {}
        done = []
        t0 = clock()
        max_conc = 0
        replays = 0
        advance_s = 0.0

        def prepare_window():
            # the first `prepare_ahead` queued requests have their states under way
            for k in itertools.islice(queue, prepare_ahead):
                if k not in states:
                    states[k] = prep.submit(requests[k].state)

        while pending or queue or active:
            now = clock() - t0
            while pending and (not realtime or requests[pending[0]].arrival <= now):
                k = pending.popleft()
                req = requests[k]
                if instance_bytes(self.precision, req.n, req.n) > self.hbm_budget:
                    raise ValueError(f"request {k} ({req.n}^2) exceeds the whole HBM budget")
                queue.append(k)
            prepare_window()
            while queue:
                k = queue[0]
                req = requests[k]
                # the head is always inside the window, whose edge moves with every admission
                fut = states[k]
                if active and not fut.done():
                    break  # the device has work; admit this one after the next replay
                try:
                    iid = self.engine.add(fut.result(), req.t_end, cfl=self.cfl,
                                          gamma=req.gamma, glm_alpha=self.glm_alpha, bc=req.bc,
                                          max_steps=req.max_steps, **_scheme(req))
                except CapacityFull:
                    if not active:
                        # nothing resident and still no room: it can never be admitted
                        raise ValueError(f"request {k} ({req.n}^2) exceeds the whole HBM "
                                         "budget") from None
                    break
                queue.popleft()
                del states[k]
                active[iid] = (k, clock() - t0)
                prepare_window()
            max_conc = max(max_conc, len(active))
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, requests[pending[0]].arrival - (clock() - t0)))
                continue
            t_adv = clock()
            self.engine.advance(max_replays=1)
            advance_s += clock() - t_adv
            replays += 1
            for iid in self.engine.poll_finished():
                k, adm = active.pop(iid)
                req = requests[k]
                try:
                    res, _ = self.engine.result(iid)
                    steps, t, err = res.steps, res.t, None
                except (ValueError, RuntimeError) as e:
                    # the request's own solve failed (dt <= 0, max_steps before t_end): report
                    # it with the request and keep serving the others
                    steps, t, err = 0, float("nan"), str(e)
                state = self.engine.download(iid) if keep_states and err is None else None
                self.engine.remove(iid)
                done.append(Completed(k, req.n, req.ic, req.arrival if realtime else 0.0, adm,
                                      clock() - t0, steps, t, state, err))
        return ServeReport(done, clock() - t0, max_conc, replays, advance_s)
