"""The serving leg's host paths under ThreadSanitizer and AddressSanitizer (CPU): initial
states built on several threads at once while other threads pack and unpack large states
through the shared copy pool (tests/host/host_threads_check.cpp, g++ -fsanitize=...). Any
sanitizer report fails the test, as does a result that differs from the same work on one
thread."""
import os
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
CSRC = ROOT / "mhd_gpu_solver_amd" / "csrc"
HIP_INCLUDE = pathlib.Path("/opt/rocm/include")


@pytest.mark.parametrize("sanitizer", ["thread", "address"])
def test_host_paths_under_sanitizer(tmp_path, sanitizer):
    gxx = shutil.which("g++")
    if gxx is None or not HIP_INCLUDE.exists():
        pytest.skip("g++ or the HIP headers are not available")
    exe = tmp_path / f"host_threads_{sanitizer}"
    cmd = [gxx, "-std=c++17", "-O1", "-g", f"-fsanitize={sanitizer}", "-pthread",
           "-D__HIP_PLATFORM_AMD__", f"-I{HIP_INCLUDE}", f"-I{CSRC}", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "host" / "host_threads_check.cpp"), str(CSRC / "mhd_host.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66",
               ASAN_OPTIONS="halt_on_error=1 exitcode=66 detect_leaks=1")
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    print(res.stdout[-2000:], res.stderr[-4000:])
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ThreadSanitizer" not in res.stderr and "AddressSanitizer" not in res.stderr
    assert res.stdout.strip().endswith("ok")
