"""bench.py's roofline block on the CPU: algorithmic bytes of the dominant stage kernel over
its mean launch time, against the MI355X HBM peak, with the measured HBM traffic taken from
the committed FETCH_SIZE / WRITE_SIZE summaries; and the refusal of more than one GPU."""
import importlib.util
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent.parent


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_roofline_fields_and_arithmetic():
    b = load_bench()
    cells = 13334528
    # stage 1: 100 launches, 35 ms in all; stage 2: 100 launches, 40 ms: stage 2 dominates
    r = b.roofline([(35.0, 100), (40.0, 100)], cells, 0.75, 4, "k_batch_stage",
                   traffic=lambda dom: b.measured_traffic(f"k_batch_stage<{dom + 1}", 4,
                                                          "ensemble_512", "k_batch_stage"))
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["kernel"] == "k_batch_stage<2>"
    want = 27 * 4 * cells / 0.4e-3 / 1e9        # 27 fp32 values per cell in 0.4 ms
    assert abs(r["achieved"] - want) < 0.1
    assert abs(r["frac"] - want / 8000.0) < 1e-4
    # traffic from the committed summary: within 1.5x of the algorithmic bytes per launch
    assert r["traffic"] is not None and 1.0 < r["traffic"] / (27 * 4 * cells) < 1.5


def test_measured_traffic_absent_for_unknown_workload():
    b = load_bench()
    assert b.measured_traffic("k_stage<2", 8, "no_such_workload") is None


def test_more_than_one_gpu_refused():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2 and "one GPU" in res.stderr
