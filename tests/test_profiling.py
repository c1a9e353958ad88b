"""The rocprofv3 CSV summaries (mhd_gpu_solver_amd/profiling.py) on small synthetic traces."""
import csv

from mhd_gpu_solver_amd import profiling


def test_short_names_keep_template_arguments():
    assert profiling._short(
        "void mhdhip::k_batch_stage<2, 0, 0, 1, float, 0>(mhdhip::BatchTables, int)"
    ) == "k_batch_stage<2, 0, 0, 1, float, 0>"
    # kernels in an unnamed namespace: the parenthesis of "(anonymous namespace)" is not the
    # start of the argument list
    assert profiling._short(
        "void mhdhip::(anonymous namespace)::k_stream2<double>(double const*, double*, long long)"
    ) == "k_stream2<double>"
    assert profiling._short("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_kernel_table_splits_active_and_noop_launches(tmp_path):
    path = tmp_path / "trace.csv"
    name = "void mhdhip::k_stage<2, 0, 0, 1, float, 0>(mhdhip::StageBuffers, int)"
    anon = "void mhdhip::(anonymous namespace)::k_stream1<float>(float const*, float*, long long)"
    # durations in ns: three active launches of 400, 500, 600 us, one no-op of 4 us
    launches = [(name, 400_000), (name, 4_000), (name, 500_000), (name, 600_000),
                (anon, 200_000)]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        t = 1_000
        for k, d in launches:
            w.writerow({"Kernel_Name": k, "Start_Timestamp": t, "End_Timestamp": t + d})
            t += d + 1_000
    table = profiling.kernel_table(str(path), last=2).splitlines()
    rows = {r.split("|")[1].strip(): [c.strip() for c in r.split("|")[2:-1]] for r in table[2:]}
    assert rows["k_stage<2, 0, 0, 1, float, 0>"] == ["4", "3", "500.0", "400.0", "600.0",
                                                      "4.00", "550.0"]
    assert rows["k_stream1<float>"][:3] == ["1", "1", "200.0"]
