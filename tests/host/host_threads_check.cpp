// Host paths of the serving leg under a thread sanitizer or address sanitizer build: request
// states built by mhdhip::init_ic on several threads at once (as the serving loop's
// preparation pool does through mhd_hip_init_ic) while other threads pack and unpack large
// states through the shared CopyPool (gather_planes / scatter_planes, as batch admission and
// download do). Every result must equal the same work done on one thread. Built and run by
// tests/test_host_threads.py; prints "ok" and exits 0 when every check holds.
#include "mhd_host.hpp"
#include "mhd_transfer.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

struct State {
    int nx, ny;
    std::vector<double> data;  // 9 planes of (ny + 4) (nx + 4)
    double* plane(int c) { return data.data() + static_cast<size_t>(c) * (nx + 4) * (ny + 4); }
};

State make_ic(const char* name, int n, uint64_t seed)
{
    State s{n, n, std::vector<double>(static_cast<size_t>(9) * (n + 4) * (n + 4), 0.0)};
    mhd_hip_grid g{n, n, 2, 0.0, 1.0, 0.0, 1.0};
    double* comps[9];
    for (int c = 0; c < 9; ++c) comps[c] = s.plane(c);
    mhdhip::init_ic(name, g, 5.0 / 3.0, 1.0, seed, comps);
    return s;
}

bool same(const std::vector<double>& a, const std::vector<double>& b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0;
}

}  // namespace

int main()
{
    struct Job {
        const char* ic;
        int n;
        uint64_t seed;
    };
    const std::vector<Job> jobs = {{"kh_tanh", 300, 1},   {"rotor", 257, 0},
                                   {"alfven_oblique", 190, 0}, {"kh_tanh", 211, 99},
                                   {"orszag_tang_b0", 160, 0}, {"rotor_b0", 128, 0}};
    std::vector<State> serial;
    for (const Job& j : jobs) serial.push_back(make_ic(j.ic, j.n, j.seed));

    // large states for the copy pool: 9 planes of 1.2M elements (above the pool's threshold)
    const size_t n = 1200 * 1000;
    std::vector<double> src(9 * n);
    for (size_t k = 0; k < src.size(); ++k) src[k] = std::sin(0.001 * static_cast<double>(k));

    int failures = 0;
    std::vector<std::thread> threads;
    std::vector<std::vector<State>> built(3);
    for (int t = 0; t < 3; ++t) {
        threads.emplace_back([&, t] {
            for (int rep = 0; rep < 2; ++rep) {
                for (size_t k = t; k < jobs.size(); k += 3) {
                    built[t].push_back(make_ic(jobs[k].ic, jobs[k].n, jobs[k].seed));
                }
            }
        });
    }
    std::vector<int> copy_ok(2, 1);
    for (int t = 0; t < 2; ++t) {
        threads.emplace_back([&, t] {
            std::vector<float> packed(9 * n);
            std::vector<double> back(9 * n, 0.0);
            const double* in[9];
            double* out[9];
            for (int c = 0; c < 9; ++c) {
                in[c] = src.data() + c * n;
                out[c] = back.data() + c * n;
            }
            for (int rep = 0; rep < 6; ++rep) {
                mhdhip::gather_planes(packed.data(), in, 9, n);
                mhdhip::scatter_planes(out, packed.data(), 9, n);
                for (size_t k = 0; k < back.size(); k += 997) {
                    if (back[k] != static_cast<double>(static_cast<float>(src[k]))) {
                        copy_ok[t] = 0;
                    }
                }
            }
        });
    }
    for (auto& th : threads) th.join();
    for (int t = 0; t < 3; ++t) {
        size_t idx = 0;
        for (int rep = 0; rep < 2; ++rep) {
            for (size_t k = t; k < jobs.size(); k += 3, ++idx) {
                if (!same(built[t][idx].data, serial[k].data)) {
                    std::printf("FAIL: %s %d built on thread %d differs\n", jobs[k].ic, jobs[k].n, t);
                    ++failures;
                }
            }
        }
    }
    for (int t = 0; t < 2; ++t) {
        if (!copy_ok[t]) {
            std::printf("FAIL: copy pool round trip on thread %d\n", t);
            ++failures;
        }
    }
    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
