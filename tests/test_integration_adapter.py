"""The C++ adapter documented for maintainers of the reference (HipSolver : DeviceSolver over the
C ABI, section 2 of the integration notes at the repository root) compiles against the
reference's real DeviceSolver interface and this backend's include/mhd_hip.h. Needs the
reference sources, which exist only where they were unpacked; the test skips elsewhere."""
import pathlib
import re
import shutil
import subprocess

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
