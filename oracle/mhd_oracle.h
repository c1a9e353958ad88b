/*
 * CPU fp64 oracle for the 2-D GLM-MHD MUSCL/HLL SSP-RK2 step.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (mhd_gpu_solver_amd/, include/) may link,
 * import or call this library; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark use it, and only as the checker.
 *
 * State layout (same as the reference StateField and the device layout): nine component planes
 * (rho, mx, my, mz, energy, bx, by, bz, psi), each (ny + 2 ng) x (nx + 2 ng) doubles, row-major,
 * concatenated: U[c * n + j * nx_tot + i], n = nx_tot * ny_tot.
 */
#ifndef MHD_ORACLE_H
#define MHD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Enum values follow the reference's enum order (BoundaryConditionType, SlopeLimiter). */
enum { ORC_PERIODIC = 0, ORC_REFLECTIVE = 1, ORC_OUTFLOW = 2 };
enum { ORC_MINMOD = 0, ORC_MC = 1 };

/* Return codes. */
enum { ORC_OK = 0, ORC_EINVAL = 1, ORC_EMAXSTEPS = 2 };

typedef struct {
    int nx, ny, ng;
    double x_min, x_max, y_min, y_max;
} orc_grid;

typedef struct {
    double t_end;
    double cfl;
    double gamma;
    double glm_alpha;
    int limiter;
    int ohmic, hall, ambipolar;
    double eta_ohm, eta_hall, eta_ambipolar;
    int max_steps;
} orc_params;

/* Fills p with the reference defaults (cfl 0.4, gamma 5/3, alpha 0.1, MC, ideal, 1e6 steps). */
void orc_default_params(orc_params* p);

/* Limiter on consecutive differences a, b. */
double orc_limit(int limiter, double a, double b);

/* One face flux in the rotated (n, t) frame. Each W is (rho, vn, vt, vz, p, bn, bt, bz, psi);
 * F receives (rho, mn, mt, mz, E, bn, bt, bz, psi) fluxes along n. */
void orc_face_flux(int limiter, const double* Wmm, const double* Wm, const double* W0,
                   const double* Wp, double gamma, double inv_gm1, double c_h, double* F);

/* Fast magnetosonic speed along n for a rotated primitive state W[9]. */
double orc_fast_speed(const double* W, double gamma);

/* max(|vx| + cf_x, |vy| + cf_y) of one cell from its conserved state. */
double orc_cell_max_speed(const double* U9, double gamma);

/* Non-ideal electric field E[3] from face J[3], B[3]. */
void orc_nonideal_e(const double* J, const double* B, const orc_params* p, double* E);

/* Periodic ghost source index (reference periodic_source). */
int orc_periodic_source(int g, int n, int ng, int low);

/* Ghost fill, x pass then y pass. Returns ORC_EINVAL for Reflective. */
int orc_apply_bc(const orc_grid* g, double* U, int bc_x, int bc_y);

/* Reflective slip fill of solid cells (mask: ny * nx bytes, 1 = fluid). */
void orc_wall_fill(const orc_grid* g, double* U, const uint8_t* mask, double gamma);

/* Spatial operator L(U) into rhs (interior only; ghosts of rhs untouched). */
void orc_rhs(const orc_grid* g, const double* U, double* rhs, const orc_params* p, double c_h,
             const uint8_t* mask);

/* CFL: returns dt, writes c_h. mask may be NULL. */
double orc_cfl(const orc_grid* g, const double* U, const orc_params* p, const uint8_t* mask,
               double* c_h);

/* Time integration with the reference solve() contract. If allow_max_steps is nonzero, stopping
 * at max_steps is not an error (fixed step-count runs). dt_log (nullable) receives the first
 * dt_cap dt values. */
int orc_solve(const orc_grid* g, int bc_x, int bc_y, const uint8_t* mask, int reflective_walls,
              const orc_params* p, int allow_max_steps, double* U, double* dt_log, int dt_cap,
              double* t_out, int* steps_out, double* ch_out);

/* orc_solve with exactly `nsteps` steps of the given dt sequence (c_h from the CFL reduction). */
int orc_solve_dt(const orc_grid* g, int bc_x, int bc_y, const uint8_t* mask, int reflective_walls,
                 const orc_params* p, double* U, const double* dts, int nsteps, double* t_out,
                 int* steps_out, double* ch_out);

/* max |dBx/dx + dBy/dy| over interior cells by central differences of U as stored (ghosts
 * included), max by std::max rule. Restates inititial_condition.hpp:273-287. */
double orc_max_div_b(const orc_grid* g, const double* U);

/* Reference-named initial condition on the interior (conserved), then ghost fill.
 * Returns ORC_EINVAL for an unknown name. */
int orc_init_ic(const char* name, const orc_grid* g, double gamma, double* U);

/* Number of OpenMP threads the oracle uses (1 if built without OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
