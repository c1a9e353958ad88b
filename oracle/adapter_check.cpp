// Checker (test infrastructure): the documented C++ adapter (mhd_gpu_solver_amd/cpp/
// hip_solver.cpp, HipSolver : mhd::DeviceSolver over include/mhd_hip.h) linked against
// libmhd_hip.so and the reference's own headers, driven the way the reference's MHDSolver
// drives a device backend (include/mhd/solver_api.hpp:294-321: configure, upload, solve,
// download).
//   * without a device: hip_available() is false and make_hip_solver throws
//     std::runtime_error;
//   * with one: fp64 reproduces the reference's CPU solve of the same Orszag-Tang state bit
//     for bit (state, t, steps); fp32 stays within 1e-4 relative L1 of it; t_end < 0 throws
//     std::invalid_argument and max_steps before t_end std::runtime_error
//     (tests/test_gpu_backends.cpp:126-157).
// Prints "adapter ok" and exits 0 when every check holds.
#include "hip_solver.hpp"
#include "mhd/solver_api.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

int failures = 0;

void fail(const char* what)
{
    std::printf("FAIL: %s\n", what);
    ++failures;
}

std::vector<const std::vector<double>*> comps(const StateField& U)
{
    return {&U.rho, &U.mx, &U.my, &U.mz, &U.energy, &U.bx, &U.by, &U.bz, &U.psi};
}

bool bit_equal(const StateField& a, const StateField& b)
{
    const auto ca = comps(a), cb = comps(b);
    for (size_t c = 0; c < ca.size(); ++c) {
        if (ca[c]->size() != cb[c]->size() ||
            std::memcmp(ca[c]->data(), cb[c]->data(), ca[c]->size() * sizeof(double)) != 0) {
            return false;
        }
    }
    return true;
}

double rel_l1(const std::vector<double>& a, const std::vector<double>& b)
{
    double num = 0.0, den = 0.0;
    for (size_t k = 0; k < a.size(); ++k) {
        num += std::fabs(a[k] - b[k]);
        den += std::fabs(b[k]);
    }
    return den > 0.0 ? num / den : num;
}

}  // namespace

int main()
{
    const bool avail = mhd::hip_available();
    std::printf("hip_available %d\n", avail ? 1 : 0);
    if (!avail) {
        try {
            auto dev = mhd::make_hip_solver(Grid2D(16, 16, 0.0, 1.0, 0.0, 1.0, 2),
                                            mhd::Precision::FP64);
            fail("a solver was created without a device");
        } catch (const std::runtime_error& e) {
            std::printf("no device: %s\n", e.what());
        }
    } else {
        MHDSolver ref(64, 48);
        ref.set_limiter_mc();
        ref.initialize("orszag_tang");
        const StateField init = ref.state();
        const SolveResult rc = ref.run(0.02);
        for (const mhd::Precision prec : {mhd::Precision::FP64, mhd::Precision::FP32}) {
            auto dev = mhd::make_hip_solver(ref.grid(), prec);
            if (dev->precision() != prec) fail("precision()");
            if (dev->device_name().rfind("HIP: ", 0) != 0) fail("device_name()");
            dev->configure(ref.boundary_conditions(), nullptr, true);
            dev->upload(init);
            SolveParams p = ref.params();
            p.t_end = 0.02;
            const SolveResult r = dev->solve(p);
            StateField out = init;
            dev->download(out);
            if (prec == mhd::Precision::FP64) {
                if (r.steps != rc.steps || r.t != rc.t) fail("fp64 steps / t");
                if (!bit_equal(out, ref.state())) fail("fp64 state differs from the CPU solve");
            } else {
                if (std::abs(r.steps - rc.steps) > 1) fail("fp32 steps");
                if (rel_l1(out.rho, ref.state().rho) > 1e-4) fail("fp32 rho L1");
                if (rel_l1(out.energy, ref.state().energy) > 1e-4) fail("fp32 energy L1");
            }
            p.t_end = -1.0;
            try {
                dev->solve(p);
                fail("t_end < 0 accepted");
            } catch (const std::invalid_argument&) {
            }
            p.t_end = 1.0;
            p.max_steps = 3;
            try {
                dev->solve(p);
                fail("max_steps not enforced");
            } catch (const std::invalid_argument&) {
                fail("max_steps raised invalid_argument");
            } catch (const std::runtime_error&) {
            }
        }
        std::printf("steps %d t %.17g\n", rc.steps, rc.t);
    }
    if (failures) return 1;
    std::printf("adapter ok\n");
    return 0;
}
