"""Seeded workload generators (host only)."""

import numpy as np

from mhd_gpu_solver_amd import workloads


def test_ensemble_is_seeded_and_in_range():
    a = workloads.ensemble(512, seed=2024)
    b = workloads.ensemble(512, seed=2024)
    assert [(s.ic, s.n, s.gamma, s.b0) for s in a] == [(s.ic, s.n, s.gamma, s.b0) for s in b]
    assert {s.n for s in a} == {64, 96, 128, 192, 256}
    assert all(1.4 <= s.gamma <= 5 / 3 for s in a)
    ot = [s for s in a if s.ic == "orszag_tang_b0"]
    ro = [s for s in a if s.ic == "rotor_b0"]
    assert 200 < len(ot) < 312 and len(ot) + len(ro) == 512
    assert all(0.5 * workloads.B0_STD <= s.b0 <= 2 * workloads.B0_STD for s in ot)
    assert all(0.5 <= s.b0 <= 2.0 for s in ro)
    assert all(s.bc == (0, 0) for s in ot) and all(s.bc == (2, 2) for s in ro)


def test_request_stream_poisson_and_ranges():
    r = workloads.request_stream(4000, rate=200.0, seed=7)
    t = np.array([s.arrival for s in r])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1 / 200) < 0.1 / 200
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1          # exponential: cv = 1
    assert min(s.n for s in r) >= 64 and max(s.n for s in r) <= 1024
    assert all(0.05 <= s.t_end <= 0.5 for s in r)
    kinds = {s.ic for s in r}
    assert kinds == {"alfven_oblique", "kh_tanh", "rotor"}
    assert len({s.seed for s in r if s.ic == "kh_tanh"}) > 100
