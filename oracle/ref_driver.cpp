// C-ABI driver over the reference's own header-only CPU solver. TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile against the reference headers where they lie (include/ of the
// reference checkout); the output goes to oracle/_ref/ and never into the repository history.
// Used by tests to pin the C oracle (bit-exact when both are built with -ffp-contract=off) and,
// in a Release-flag build, as the timed CPU baseline.
//
// Entry points mirror oracle/mhd_oracle.h so tests can call either with the same arguments.

#include "boundary_conditions/boundary_conditions.hpp"
#include "grid/grid.hpp"
#include "initial_condition/inititial_condition.hpp"
#include "mhd/cfl.hpp"
#include "mhd/kernels/muscl_hll.hpp"
#include "mhd/mask.hpp"
#include "mhd/nonideal.hpp"
#include "mhd/parallel.hpp"
#include "mhd/rhs.hpp"
#include "mhd/simd.hpp"
#include "mhd/time_integrator.hpp"
#include "state/state_field.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <vector>

namespace {

struct RefParams {
    double t_end;
    double cfl;
    double gamma;
    double glm_alpha;
    int limiter;
    int ohmic, hall, ambipolar;
    double eta_ohm, eta_hall, eta_ambipolar;
    int max_steps;
};

Grid2D make_grid(int nx, int ny, int ng, double x0, double x1, double y0, double y1)
{
    return Grid2D(nx, ny, x0, x1, y0, y1, ng);
}

BoundaryConditionType bc_of(int v)
{
    switch (v) {
    case 0:
        return BoundaryConditionType::Periodic;
    case 1:
        return BoundaryConditionType::Reflective;
    default:
        return BoundaryConditionType::Outflow;
    }
}

SolveParams params_of(const RefParams* p)
{
    SolveParams s{};
    s.t_end = p->t_end;
    s.cfl = p->cfl;
    s.gamma = p->gamma;
    s.glm_alpha = p->glm_alpha;
    s.limiter = p->limiter == 0 ? SlopeLimiter::Minmod : SlopeLimiter::MC;
    s.nonideal = NonIdealConfig::ideal();
    s.nonideal.ohmic = p->ohmic != 0;
    s.nonideal.hall = p->hall != 0;
    s.nonideal.ambipolar = p->ambipolar != 0;
    s.nonideal.eta_ohm = p->eta_ohm;
    s.nonideal.eta_hall = p->eta_hall;
    s.nonideal.eta_ambipolar = p->eta_ambipolar;
    s.max_steps = p->max_steps;
    return s;
}

std::vector<double>* plane(StateField& U, int c)
{
    std::vector<double>* planes[9] = {&U.rho, &U.mx, &U.my, &U.mz, &U.energy,
                                      &U.bx,  &U.by, &U.bz, &U.psi};
    return planes[c];
}

void load_state(StateField& U, const double* src)
{
    const std::size_t n = U.rho.size();
    for (int c = 0; c < 9; ++c) {
        std::memcpy(plane(U, c)->data(), src + c * n, n * sizeof(double));
    }
}

void store_state(StateField& U, double* dst)
{
    const std::size_t n = U.rho.size();
    for (int c = 0; c < 9; ++c) {
        std::memcpy(dst + c * n, plane(U, c)->data(), n * sizeof(double));
    }
}

FluidMask make_mask(const std::uint8_t* mask, int nx, int ny)
{
    FluidMask m(nx, ny, true);
    if (mask != nullptr) {
        m.set_from_interior(mask, static_cast<std::size_t>(nx) * ny, nx, ny);
    }
    return m;
}

} // namespace

extern "C" {

void ref_set_threads(int n)
{
    mhd::set_num_threads(n);
}

int ref_num_threads()
{
    return mhd::num_threads();
}

const char* ref_simd_backend()
{
    return mhd::simd::kBackendName;
}

// Same contract as orc_solve. Without a dt log the reference's own solve() runs; with one, the
// loop is driven step by step through the reference's compute_cfl_dt / ssp_rk2_step.
int ref_solve(int nx, int ny, int ng, double x0, double x1, double y0, double y1, int bc_x,
              int bc_y, const std::uint8_t* mask, int walls, const RefParams* rp,
              int allow_max_steps, double* Uio, double* dt_log, int dt_cap, double* t_out,
              int* steps_out, double* ch_out)
{
    try {
        const Grid2D grid = make_grid(nx, ny, ng, x0, x1, y0, y1);
        const BoundaryConditions bc(bc_of(bc_x), bc_of(bc_y));
        StateField U(grid.get_size_x(), grid.get_size_y());
        load_state(U, Uio);
        TimeIntegratorWorkspace work(grid);
        const FluidMask fm = make_mask(mask, nx, ny);
        const FluidMask* mp = fm.all_fluid() ? nullptr : &fm;
        const SolveParams params = params_of(rp);
        SolveResult r{};
        if (dt_log == nullptr && !allow_max_steps) {
            r = solve(U, work, grid, bc, params, mp, walls != 0);
        } else {
            while (r.t < params.t_end && r.steps < params.max_steps) {
                const CFLResult cfl =
                    compute_cfl_dt(U, grid, params.gamma, params.cfl, params.nonideal, mp);
                r.c_h = cfl.c_h;
                const double dt = std::min(cfl.dt, params.t_end - r.t);
                if (dt_log != nullptr && r.steps < dt_cap) {
                    dt_log[r.steps] = dt;
                }
                ssp_rk2_step(U, work, grid, bc, params.gamma, dt, cfl.c_h, params.glm_alpha,
                             params.limiter, params.nonideal, mp, walls != 0);
                r.t += dt;
                ++r.steps;
            }
            if (r.t < params.t_end && r.steps >= params.max_steps && !allow_max_steps) {
                return 2;
            }
            apply_boundary_conditions(U, grid, bc);
            if (mp != nullptr && walls != 0) {
                apply_solid_wall_fill(U, grid, *mp, params.gamma);
            }
        }
        store_state(U, Uio);
        *t_out = r.t;
        *steps_out = r.steps;
        *ch_out = r.c_h;
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (const std::runtime_error&) {
        return 2;
    } catch (...) {
        return 3;
    }
}

int ref_init_ic(const char* name, int nx, int ny, int ng, double x0, double x1, double y0,
                double y1, double gamma, double* Uio)
{
    try {
        const Grid2D grid = make_grid(nx, ny, ng, x0, x1, y0, y1);
        StateField U(grid.get_size_x(), grid.get_size_y());
        load_state(U, Uio);
        initialize_state_field(U, grid, parse_initial_condition_name(name), gamma);
        store_state(U, Uio);
        return 0;
    } catch (...) {
        return 1;
    }
}

int ref_apply_bc(int nx, int ny, int ng, int bc_x, int bc_y, double* Uio)
{
    try {
        const Grid2D grid = make_grid(nx, ny, ng, 0.0, 1.0, 0.0, 1.0);
        StateField U(grid.get_size_x(), grid.get_size_y());
        load_state(U, Uio);
        apply_boundary_conditions(U, grid, BoundaryConditions(bc_of(bc_x), bc_of(bc_y)));
        store_state(U, Uio);
        return 0;
    } catch (...) {
        return 1;
    }
}

int ref_wall_fill(int nx, int ny, int ng, const std::uint8_t* mask, double gamma, double* Uio)
{
    try {
        const Grid2D grid = make_grid(nx, ny, ng, 0.0, 1.0, 0.0, 1.0);
        StateField U(grid.get_size_x(), grid.get_size_y());
        load_state(U, Uio);
        const FluidMask fm = make_mask(mask, nx, ny);
        apply_solid_wall_fill(U, grid, fm, gamma);
        store_state(U, Uio);
        return 0;
    } catch (...) {
        return 1;
    }
}

int ref_rhs(int nx, int ny, int ng, double x0, double x1, double y0, double y1, const double* Uin,
            double* rhs_out, const RefParams* rp, double c_h, const std::uint8_t* mask)
{
    try {
        const Grid2D grid = make_grid(nx, ny, ng, x0, x1, y0, y1);
        StateField U(grid.get_size_x(), grid.get_size_y());
        load_state(U, Uin);
        StateField R(grid.get_size_x(), grid.get_size_y());
        load_state(R, rhs_out);
        RHSWorkspace work(grid.get_size_x(), grid.get_size_y());
        const FluidMask fm = make_mask(mask, nx, ny);
        const FluidMask* mp = fm.all_fluid() ? nullptr : &fm;
        const SolveParams params = params_of(rp);
        compute_rhs(U, R, grid, work, params.gamma, c_h, params.glm_alpha, params.limiter,
                    params.nonideal, mp);
        store_state(R, rhs_out);
        return 0;
    } catch (...) {
        return 1;
    }
}

double ref_cfl(int nx, int ny, int ng, double x0, double x1, double y0, double y1,
               const double* Uin, const RefParams* rp, const std::uint8_t* mask, double* c_h)
{
    const Grid2D grid = make_grid(nx, ny, ng, x0, x1, y0, y1);
    StateField U(grid.get_size_x(), grid.get_size_y());
    load_state(U, Uin);
    const FluidMask fm = make_mask(mask, nx, ny);
    const FluidMask* mp = fm.all_fluid() ? nullptr : &fm;
    const SolveParams params = params_of(rp);
    const CFLResult r = compute_cfl_dt(U, grid, params.gamma, params.cfl, params.nonideal, mp);
    *c_h = r.c_h;
    return r.dt;
}

double ref_max_div_b(int nx, int ny, int ng, double x0, double x1, double y0, double y1,
                     const double* Uin)
{
    const Grid2D grid = make_grid(nx, ny, ng, x0, x1, y0, y1);
    StateField U(grid.get_size_x(), grid.get_size_y());
    load_state(U, Uin);
    return max_abs_div_b(U, grid);
}

void ref_face_flux(int limiter, const double* Wmm, const double* Wm, const double* W0,
                   const double* Wp, double gamma, double inv_gm1, double c_h, double* F)
{
    using mhd::kernels::PrimN;
    auto mk = [](const double* w) {
        PrimN<double> p{w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], w[8]};
        return p;
    };
    const PrimN<double> a = mk(Wmm), b = mk(Wm), c = mk(W0), d = mk(Wp);
    mhd::kernels::FluxN<double> f{};
    if (limiter == 0) {
        f = mhd::kernels::muscl_hll_face<SlopeLimiter::Minmod>(a, b, c, d, gamma, inv_gm1, c_h);
    } else {
        f = mhd::kernels::muscl_hll_face<SlopeLimiter::MC>(a, b, c, d, gamma, inv_gm1, c_h);
    }
    const double out[9] = {f.rho, f.mn, f.mt, f.mz, f.e, f.bn, f.bt, f.bz, f.psi};
    std::memcpy(F, out, sizeof(out));
}

double ref_cell_max_speed(const double* U9, double gamma)
{
    return mhd::kernels::cell_max_signal_speed<double>(U9[0], U9[1], U9[2], U9[3], U9[4], U9[5],
                                                       U9[6], U9[7], gamma);
}

void ref_nonideal_e(const double* J, const double* B, const RefParams* rp, double* E)
{
    const SolveParams params = params_of(rp);
    const Vec3 e = nonideal_electric_field(Vec3{J[0], J[1], J[2]}, Vec3{B[0], B[1], B[2]},
                                           params.nonideal);
    E[0] = e.x;
    E[1] = e.y;
    E[2] = e.z;
}

// Timed CPU baseline: `steps` SSP-RK2 steps of the state in Uio (periodic or outflow), using the
// reference's own solve() with t_end large and max_steps = steps. Returns wall seconds.
double ref_time_steps(int nx, int ny, int ng, double x0, double x1, double y0, double y1,
                      int bc_x, int bc_y, const RefParams* rp, int steps, double* Uio)
{
    const Grid2D grid = make_grid(nx, ny, ng, x0, x1, y0, y1);
    const BoundaryConditions bc(bc_of(bc_x), bc_of(bc_y));
    StateField U(grid.get_size_x(), grid.get_size_y());
    load_state(U, Uio);
    TimeIntegratorWorkspace work(grid);
    SolveParams params = params_of(rp);
    params.t_end = 1.0e30;
    params.max_steps = steps;
    const auto t0 = std::chrono::steady_clock::now();
    try {
        solve(U, work, grid, bc, params, nullptr, true);
    } catch (const std::runtime_error&) {
        // max_steps reached before t_end: expected for a fixed step count
    }
    const auto t1 = std::chrono::steady_clock::now();
    store_state(U, Uio);
    return std::chrono::duration<double>(t1 - t0).count();
}

} // extern "C"
