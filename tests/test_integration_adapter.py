"""The C++ adapter for maintainers of the reference (HipSolver : DeviceSolver over the C ABI,
mhd_gpu_solver_amd/cpp/hip_solver.cpp, shown in section 2 of the integration notes at the
repository root):
* the notes show the committed source;
* it compiles against the reference's real DeviceSolver interface and include/mhd_hip.h;
* linked with libmhd_hip.so and the reference's headers (oracle/_ref/hip_adapter_check, built
  by `make -C oracle adapter`), it reports no device and maps the library's ENODEV to
  std::runtime_error on a machine without a GPU; tests/test_gpu_adapter.py runs it on the GPU.
The compile check needs the reference sources, which exist only where they were unpacked."""
import pathlib
import re
import shutil
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
REF_INCLUDE = pathlib.Path("/root/reference/include")

# the two declarations section 1 of the integration notes adds to device_solver.hpp
PRELUDE = """#include <memory>
#include "mhd/gpu/device_solver.hpp"
namespace mhd {
bool hip_available();
std::unique_ptr<DeviceSolver> make_hip_solver(const Grid2D& grid, Precision precision);
}
"""


def adapter_source() -> str:
    text = (ROOT / "INTEGRATION.md").read_text()
    sec = text[text.index("## 2. `src/hip/hip_solver.cpp`"):text.index("## 3.")]
    return re.search(r"```cpp\n(.*?)```", sec, re.S).group(1)


@pytest.mark.skipif(not (REF_INCLUDE / "mhd/gpu/device_solver.hpp").exists(),
                    reason="reference sources not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_documented_adapter_compiles_against_reference(tmp_path):
    src = tmp_path / "hip_solver.cpp"
    src.write_text(PRELUDE + adapter_source())
    res = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                          f"-I{REF_INCLUDE}", f"-I{ROOT / 'include'}", str(src)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_notes_show_the_committed_adapter():
    src = (ROOT / "mhd_gpu_solver_amd" / "cpp" / "hip_solver.cpp").read_text()
    body = src[src.index('#include "hip_solver.hpp"'):]
    shown = adapter_source().replace('#include "mhd/gpu/device_solver.hpp"',
                                     '#include "hip_solver.hpp"')
    assert body.strip() == shown.strip()


CHECK = ROOT / "oracle" / "_ref" / "hip_adapter_check"


@pytest.mark.skipif(not CHECK.exists(), reason="adapter check not built (reference build)")
def test_linked_adapter_without_device():
    sys.path.insert(0, str(ROOT))
    import mhd_gpu_solver_amd as m
    if m.hip_available():
        pytest.skip("a GPU is visible: tests/test_gpu_adapter.py covers this machine")
    res = subprocess.run([str(CHECK)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "hip_available 0" in res.stdout and "no device:" in res.stdout
    assert res.stdout.strip().endswith("adapter ok")
