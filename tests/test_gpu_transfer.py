"""Host <-> device state transfers. States of 8 MiB and more move in 16 MiB chunks through two
pinned buffers, filled and drained by host threads while the copy engine moves the other one;
smaller states use direct copies. Both must hand back exactly what went in (fp64), or its
rounding to fp32 (FP32 storage), for plane lengths that end mid-chunk."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device(hip, nx, ny, precision):
    from mhd_gpu_solver_amd import _lib
    grid = _lib.Grid(nx, ny, 2, 0.0, 1.0, 0.0, 1.0)
    return hip.DeviceSolver(grid, _lib.FP64 if precision == "fp64" else _lib.FP32)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
# 60 x 40: direct copies; 1000 x 700 and 1531 x 1023: one partial chunk per plane;
# 2101 x 2099 (4.42 M elements per plane): planes span several chunks in both precisions
@pytest.mark.parametrize("nx,ny", [(60, 40), (1000, 700), (1531, 1023), (2101, 2099)])
def test_upload_download_round_trip(hip, precision, nx, ny):
    rng = np.random.default_rng(nx * 7 + ny)
    U = rng.standard_normal((9, ny + 4, nx + 4)) * 10.0 ** rng.integers(-3, 4, (9, 1, 1))
    dev = _device(hip, nx, ny, precision)
    dev.upload(U)
    out = np.full_like(U, np.nan)
    dev.download(out)
    want = U if precision == "fp64" else U.astype(np.float32).astype(np.float64)
    assert np.array_equal(out, want)


def test_round_trips_from_threads(hip):
    """Several solvers moving large states at once share the host copy threads."""
    shapes = [(1000, 700, "fp64"), (900, 800, "fp32"), (1200, 500, "fp64"), (700, 700, "fp32")]
    rng = np.random.default_rng(5)
    states = [rng.standard_normal((9, ny + 4, nx + 4)) for nx, ny, _ in shapes]
    devs = [_device(hip, nx, ny, p) for nx, ny, p in shapes]
    outs = [np.zeros_like(U) for U in states]
    errors = []

    def work(k):
        try:
            for _ in range(3):
                devs[k].upload(states[k])
                devs[k].download(outs[k])
        except Exception as e:  # reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(shapes))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for (nx, ny, p), U, out in zip(shapes, states, outs):
        want = U if p == "fp64" else U.astype(np.float32).astype(np.float64)
        assert np.array_equal(out, want), (nx, ny, p)
