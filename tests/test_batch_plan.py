"""CPU checks of the batched engine's host planning (csrc/mhd_batch_plan.hpp): tile coverage,
one lead tile per instance, per-wave offsets, the capacity fill's balance, the timed row
choice's shortlist, and the replay accounting / no-progress check, compiled with g++ from
tests/host/batch_plan_check.cpp."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def test_batch_plan_host_checks(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "batch_plan_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    f"-I{ROOT / 'mhd_gpu_solver_amd' / 'csrc'}",
                    str(ROOT / "tests" / "host" / "batch_plan_check.cpp"), "-o", str(exe)],
                   check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.strip().endswith("ok")


def test_queue_protocol_host_model(tmp_path):
    """The dynamic tile hand-out's ticket and reset protocol, modelled on the CPU over random
    launch sequences (tests/host/queue_protocol_check.cpp), beside the earlier protocol whose
    failure it shows."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "queue_protocol_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                    f"-I{ROOT / 'mhd_gpu_solver_amd' / 'csrc'}",
                    str(ROOT / "tests" / "host" / "queue_protocol_check.cpp"), "-o", str(exe)],
                   check=True)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.strip().endswith("ok")
