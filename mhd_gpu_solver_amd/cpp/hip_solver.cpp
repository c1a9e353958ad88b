// The reference's DeviceSolver (include/mhd/gpu/device_solver.hpp:57-75) implemented over this
// backend's C ABI (include/mhd_hip.h): the adapter a maintainer adds to the reference tree as
// src/hip/hip_solver.cpp. Status codes become the reference's exception types.
#include "hip_solver.hpp"
#include "mhd_hip.h"

#include <stdexcept>
#include <string>

namespace mhd {
namespace {

void rethrow(int rc)
{
    if (rc == MHD_HIP_OK) return;
    const std::string msg = mhd_hip_last_error();
    if (rc == MHD_HIP_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

mhd_hip_params to_c(const SolveParams& p)
{
    mhd_hip_params c{};
    c.t_end = p.t_end;
    c.cfl = p.cfl;
    c.gamma = p.gamma;
    c.glm_alpha = p.glm_alpha;
    c.limiter = p.limiter == SlopeLimiter::Minmod ? MHD_HIP_MINMOD : MHD_HIP_MC;
    c.ohmic = p.nonideal.ohmic;
    c.hall = p.nonideal.hall;
    c.ambipolar = p.nonideal.ambipolar;
    c.eta_ohm = p.nonideal.eta_ohm;
    c.eta_hall = p.nonideal.eta_hall;
    c.eta_ambipolar = p.nonideal.eta_ambipolar;
    c.max_steps = p.max_steps;
    return c;
}

class HipSolver final : public DeviceSolver {
  public:
    HipSolver(const Grid2D& g, Precision prec)
    {
        const mhd_hip_grid cg{g.nx, g.ny, g.ng, g.x_min, g.x_max, g.y_min, g.y_max};
        rethrow(mhd_hip_create(&cg, prec == Precision::FP64 ? MHD_HIP_FP64 : MHD_HIP_FP32, &h_));
        n_ = static_cast<size_t>(g.get_size_x()) * g.get_size_y();
    }
    ~HipSolver() override { mhd_hip_destroy(h_); }

    std::string device_name() const override
    {
        char buf[256];
        rethrow(mhd_hip_device_name(h_, buf, sizeof(buf)));
        return buf;
    }
    Precision precision() const override
    {
        return mhd_hip_precision(h_) == MHD_HIP_FP64 ? Precision::FP64 : Precision::FP32;
    }
    void configure(const BoundaryConditions& bc, const FluidMask* mask, bool walls) override
    {
        rethrow(mhd_hip_configure(h_, static_cast<int>(bc.type_x), static_cast<int>(bc.type_y),
                                  mask ? mask->fluid.data() : nullptr, mask ? mask->size() : 0,
                                  walls ? 1 : 0));
    }
    void upload(const StateField& U) override
    {
        const double* c[9] = {U.rho.data(), U.mx.data(), U.my.data(), U.mz.data(),
                              U.energy.data(), U.bx.data(), U.by.data(), U.bz.data(),
                              U.psi.data()};
        rethrow(mhd_hip_upload(h_, c, n_));
    }
    void download(StateField& U) override
    {
        double* c[9] = {U.rho.data(), U.mx.data(), U.my.data(), U.mz.data(), U.energy.data(),
                        U.bx.data(),  U.by.data(), U.bz.data(), U.psi.data()};
        rethrow(mhd_hip_download(h_, c, n_));
    }
    SolveResult solve(const SolveParams& p) override
    {
        const mhd_hip_params c = to_c(p);
        mhd_hip_result r{};
        rethrow(mhd_hip_solve(h_, &c, &r));
        return SolveResult{r.t, r.steps, r.c_h};
    }

  private:
    mhd_hip_solver* h_ = nullptr;
    size_t n_ = 0;
};

}  // namespace

bool hip_available() { return mhd_hip_available() != 0; }

std::unique_ptr<DeviceSolver> make_hip_solver(const Grid2D& grid, Precision precision)
{
    return std::make_unique<HipSolver>(grid, precision);
}

}  // namespace mhd
