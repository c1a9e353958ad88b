"""The C++ adapter (mhd_gpu_solver_amd/cpp/hip_solver.cpp) linked with libmhd_hip.so and the
reference's own headers, on the GPU (oracle/_ref/hip_adapter_check, a reference build that
exists where the reference sources were present at build time): the reference's MHDSolver
state solved through HipSolver's configure / upload / solve / download equals the
reference's CPU solve bit for bit in fp64, stays within 1e-4 relative L1 in fp32, and the
library's status codes come back as the reference's exception types
(tests/test_gpu_backends.cpp:126-157)."""
import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu

CHECK = pathlib.Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "hip_adapter_check"


@pytest.mark.skipif(not CHECK.exists(), reason="adapter check not built (reference build)")
def test_linked_adapter_on_device(hip):
    res = subprocess.run([str(CHECK)], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "hip_available 1" in res.stdout
    assert res.stdout.strip().endswith("adapter ok")
