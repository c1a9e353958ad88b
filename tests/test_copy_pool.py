"""Host copy threads of the transfer path (mhd_gpu_solver_amd/csrc/mhd_transfer.hpp), on the
CPU: a small driver compiled with g++ against the header (HIP headers only; no HIP call is
made) checks that CopyPool::run covers [0, n) exactly once for sizes around the thread count,
also when several threads submit jobs at once, and that gather_planes / scatter_planes pack
and unpack planes with the element conversion, below and above the size where they switch to
the pool."""
import pathlib
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HIP_INCLUDE = pathlib.Path("/opt/rocm/include")

DRIVER = r"""
#include "mhd_transfer.hpp"
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <vector>

using mhdhip::CopyPool;

static int fail(const char* what, long long a, long long b)
{
    std::printf("FAIL %s %lld %lld\n", what, a, b);
    return 1;
}

static int check_cover(size_t n)
{
    std::vector<std::atomic<int>> hits(n);
    for (auto& h : hits) h = 0;
    CopyPool::instance().run(n, [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) hits[i]++;
    });
    for (size_t i = 0; i < n; ++i) {
        if (hits[i] != 1) return fail("cover", (long long)n, (long long)i);
    }
    return 0;
}

int main()
{
    for (size_t n : {0, 1, 2, 3, 7, 8, 9, 15, 16, 17, 1000, 1 << 20}) {
        if (check_cover(n)) return 1;
    }
    // several submitters at once take turns
    std::atomic<int> bad{0};
    std::vector<std::thread> ts;
    for (int t = 0; t < 6; ++t) {
        ts.emplace_back([&, t] {
            for (int rep = 0; rep < 50; ++rep) {
                if (check_cover(1000 + 37 * t + rep)) bad++;
            }
        });
    }
    for (auto& t : ts) t.join();
    if (bad) return fail("concurrent", bad, 0);
    // gather / scatter with conversion, below and above the pool threshold (1 << 20 elements)
    for (size_t n : {size_t(5), size_t(70000), size_t(1) << 17, size_t(300001)}) {
        const int count = 9;
        std::vector<std::vector<double>> planes(count, std::vector<double>(n));
        std::vector<const double*> src(count);
        for (int c = 0; c < count; ++c) {
            for (size_t k = 0; k < n; ++k) planes[c][k] = 0.1 * c + 1e-3 * k + 1.0 / 3.0;
            src[c] = planes[c].data();
        }
        std::vector<float> packed(count * n);
        mhdhip::gather_planes(packed.data(), src.data(), count, n);
        for (int c = 0; c < count; ++c) {
            for (size_t k = 0; k < n; ++k) {
                if (packed[c * n + k] != static_cast<float>(planes[c][k]))
                    return fail("gather", (long long)n, (long long)(c * n + k));
            }
        }
        std::vector<std::vector<double>> out(count, std::vector<double>(n, -1.0));
        std::vector<double*> dst(count);
        for (int c = 0; c < count; ++c) dst[c] = out[c].data();
        mhdhip::scatter_planes(dst.data(), packed.data(), count, n);
        for (int c = 0; c < count; ++c) {
            for (size_t k = 0; k < n; ++k) {
                if (out[c][k] != static_cast<double>(packed[c * n + k]))
                    return fail("scatter", (long long)n, (long long)(c * n + k));
            }
        }
    }
    std::printf("OK\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.skipif(not (HIP_INCLUDE / "hip/hip_runtime.h").exists(),
                    reason="HIP headers not present")
def test_copy_pool_and_plane_packing(tmp_path):
    src = tmp_path / "pool.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "pool"
    res = subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__",
                          f"-I{HIP_INCLUDE}", f"-I{ROOT / 'mhd_gpu_solver_amd' / 'csrc'}",
                          str(src), "-o", str(exe), "-pthread"],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.strip().endswith("OK"), run.stdout + run.stderr
