import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running case (seconds to minutes)")


@pytest.fixture(scope="session")
def oracle():
    from checkers import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from checkers import exact_reference
    ref = exact_reference()
    if ref is None:
        pytest.skip("reference build oracle/_ref/libmhd_ref.so not present")
    return ref


@pytest.fixture(scope="session")
def hip():
    """The hip backend package; fails (does not skip) when the library cannot be loaded."""
    import mhd_gpu_solver_amd as m
    if not m.hip_available():
        pytest.fail("no gfx950 device visible to the hip backend")
    return m
