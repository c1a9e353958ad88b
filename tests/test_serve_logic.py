"""Host logic of the serving loop (mhd_gpu_solver_amd/serve.py) over a fake batch engine, on
the CPU: every request completes once; admission is FIFO; initial states are built on the
preparation threads, never on the loop's thread; a request whose state is not ready waits
while others run but does not stall an idle device; a request's own solve failure is
reported with it while the others are served; the HBM-budget check rejects a request larger
than the whole budget."""
import threading
import time
from types import SimpleNamespace

import numpy as np
import pytest

from mhd_gpu_solver_amd import serve
from mhd_gpu_solver_amd.batch import CapacityFull


class FakeEngine:
    """Instances finish after a number of replays set by their t_end (in replays); at most
    `slots` are resident. A t_end < 0 marks a solve that fails."""

    def __init__(self, slots=1000, replay_s=0.0):
        self.slots = slots
        self.replay_s = replay_s
        self.replays = 0
        self.added_at = []       # replays done when each instance was admitted
        self.live = {}
        self.next_id = 0
        self.finished = []
        self.order = []          # request states in admission order
        self.add_threads = set()

    def set_schedule(self, policy, cap):
        pass

    def add(self, state, t_end, **kw):
        if len(self.live) >= self.slots:
            raise CapacityFull("full")
        self.add_threads.add(threading.get_ident())
        iid = self.next_id
        self.next_id += 1
        self.live[iid] = [int(abs(t_end)), t_end < 0, state]
        self.order.append(int(state[0, 0, 0]))
        self.added_at.append(self.replays)
        return iid

    def advance(self, max_replays=-1):
        time.sleep(self.replay_s)
        self.replays += 1
        for iid, v in self.live.items():
            if v[0] > 0:
                v[0] -= 1
                if v[0] == 0:
                    self.finished.append(iid)

    def poll_finished(self):
        out, self.finished = self.finished, []
        return out

    def result(self, iid):
        left, fails, _ = self.live[iid]
        if fails:
            raise RuntimeError("dt <= 0")
        return SimpleNamespace(steps=8, t=1.0), True

    def remove(self, iid):
        del self.live[iid]

    def download(self, iid):
        return self.live[iid][2]


class Req:
    state_threads = set()

    def __init__(self, k, replays, arrival=0.0, delay=0.0, n=16):
        self.k, self.t_end, self.arrival, self.delay, self.n = k, replays, arrival, delay, n
        self.ic, self.gamma, self.bc, self.max_steps = "fake", 5.0 / 3.0, (0, 0), 1000

    def state(self):
        Req.state_threads.add(threading.get_ident())
        time.sleep(self.delay)
        U = np.zeros((9, 4, 4))
        U[0, 0, 0] = self.k
        return U


def make_server(engine, budget=1 << 40):
    srv = object.__new__(serve.Server)
    srv.engine, srv.precision, srv.cfl, srv.glm_alpha = engine, "fp32", 0.4, 0.1
    srv.hbm_budget, srv.policy, srv.work_cap_cells = budget, "srpt", 1 << 20
    return srv


def test_all_complete_fifo_and_states_off_the_loop():
    Req.state_threads = set()
    eng = FakeEngine()
    reqs = [Req(k, replays=1 + (k * 7) % 5, delay=0.002 * (k % 3)) for k in range(24)]
    rep = make_server(eng).run(reqs, realtime=False, prepare_workers=3, prepare_ahead=4)
    assert sorted(c.index for c in rep.completed) == list(range(24))
    assert eng.order == list(range(24))                      # FIFO admission
    assert threading.get_ident() not in Req.state_threads   # states built off the loop
    assert all(c.error is None and c.latency >= 0 for c in rep.completed)
    assert rep.advance_s <= rep.makespan


class InlineExecutor:
    """Builds each state at submission, on the loop's thread: every future is already done
    when the loop looks at it, so admission runs ahead of the preparation window at once."""

    def __init__(self, max_workers=None):
        pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False

    def submit(self, fn, *args):
        from concurrent.futures import Future
        f = Future()
        f.set_result(fn(*args))
        return f


def test_admission_past_the_preparation_window(monkeypatch):
    """More admissible requests than `prepare_ahead`, every state ready at once: the loop
    admits past the first window (whose edge moves with each admission) in FIFO order."""
    monkeypatch.setattr(serve, "ThreadPoolExecutor", InlineExecutor)
    eng = FakeEngine()
    reqs = [Req(k, replays=1 + k % 3) for k in range(9)]
    rep = make_server(eng).run(reqs, realtime=False, prepare_workers=1, prepare_ahead=2)
    assert eng.order == list(range(9))
    assert eng.added_at == [0] * 9                      # all admitted before the first replay
    assert sorted(c.index for c in rep.completed) == list(range(9))


def test_admission_window_with_threads_repeated():
    """The same with real preparation threads and no delays, whatever order the futures
    finish in: every request completes once, FIFO."""
    for _ in range(20):
        eng = FakeEngine()
        reqs = [Req(k, replays=1 + k % 4) for k in range(12)]
        rep = make_server(eng).run(reqs, realtime=False, prepare_workers=3, prepare_ahead=2)
        assert eng.order == list(range(12))
        assert sorted(c.index for c in rep.completed) == list(range(12))


def test_slow_state_waits_while_others_run():
    """The head's state takes long; the loop keeps replaying the running solves meanwhile,
    and every request still completes in FIFO order."""
    eng = FakeEngine(replay_s=0.001)
    reqs = [Req(0, replays=200), Req(1, replays=2, delay=0.05), Req(2, replays=2)]
    rep = make_server(eng).run(reqs, realtime=False, prepare_workers=2, prepare_ahead=4)
    assert eng.order == [0, 1, 2]
    # request 0 kept stepping while request 1's state was being built
    assert eng.added_at[0] == 0 and eng.added_at[1] >= 5
    done = {c.index: c for c in rep.completed}
    assert done[0].admitted <= done[1].admitted <= done[2].admitted


def test_capacity_and_realtime_arrivals():
    eng = FakeEngine(slots=2)
    reqs = [Req(k, replays=3, arrival=0.002 * k) for k in range(8)]
    rep = make_server(eng).run(reqs, realtime=True, prepare_workers=2, prepare_ahead=3)
    assert sorted(c.index for c in rep.completed) == list(range(8))
    assert rep.max_concurrent <= 2
    assert all(c.finished >= c.arrival for c in rep.completed)


def test_failed_solve_reported_with_its_request():
    eng = FakeEngine()
    reqs = [Req(0, replays=2), Req(1, replays=-3), Req(2, replays=4)]
    rep = make_server(eng).run(reqs, realtime=False)
    errs = {c.index: c.error for c in rep.completed}
    assert errs[1] and errs[0] is None and errs[2] is None
    assert rep.summary()["failed"] == 1


def test_request_larger_than_budget_rejected():
    eng = FakeEngine()
    with pytest.raises(ValueError):
        make_server(eng, budget=1024).run([Req(0, replays=1, n=256)], realtime=False)


def test_request_that_never_fits_an_empty_engine_is_rejected():
    """A request the size check passes but the empty engine still cannot hold (e.g. its mask
    and wall plan on top of the state) is an error, not an endless wait."""
    eng = FakeEngine(slots=0)
    with pytest.raises(ValueError):
        make_server(eng).run([Req(0, replays=1)], realtime=False)
