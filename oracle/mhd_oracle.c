/*
 * CPU fp64 oracle: a plain-C restatement of the reference's CPU solver path for the
 * GLM-MHD MUSCL + HLL + SSP-RK2 step. TEST INFRASTRUCTURE ONLY (see mhd_oracle.h).
 *
 * Built with -ffp-contract=off so that every product and sum is rounded separately, in the
 * operation order of the reference's production CPU kernels. Each function cites the
 * reference file and lines it restates (paths relative to the reference repository root).
 *
 * Parity pins for this file live in tests/: the reference goldens and .ref diagnostics,
 * bit-exact agreement with the reference compiled from source (oracle/_ref, when built),
 * closed forms (limiters, uniform-state invariance, HLL consistency), conservation and
 * integer-tag brute force for ghost and wall-fill indexing.
 */
#include "mhd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define NC 9
/* rotated primitive slots */
enum { R_RHO = 0, R_VN, R_VT, R_VZ, R_P, R_BN, R_BT, R_BZ, R_PSI };
/* conserved / physical component planes */
enum { C_RHO = 0, C_MX, C_MY, C_MZ, C_E, C_BX, C_BY, C_BZ, C_PSI };

static const double kPi = 3.14159265358979323846;

/* scalar min/max with the reference's tie and NaN behaviour: include/mhd/simd.hpp:40-47 */
static inline double vmin(double a, double b) { return (b < a) ? b : a; }
static inline double vmax(double a, double b) { return (a < b) ? b : a; }

void orc_default_params(orc_params* p)
{
    /* include/mhd/solve_types.hpp:8-16 */
    memset(p, 0, sizeof(*p));
    p->t_end = 0.0;
    p->cfl = 0.4;
    p->gamma = 5.0 / 3.0;
    p->glm_alpha = 0.1;
    p->limiter = ORC_MC;
    p->max_steps = 1000000;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) {
        omp_set_num_threads(n);
    }
#else
    (void)n;
#endif
}

/* ---- limiters: include/mhd/kernels/muscl_hll.hpp:33-57 ---------------------------------- */
static inline double minmod(double a, double b)
{
    if (a * b <= 0.0) {
        return 0.0;
    }
    return (fabs(a) < fabs(b)) ? a : b;
}

static inline double mc_lim(double a, double b)
{
    return minmod(0.5 * (a + b), minmod(2.0 * a, 2.0 * b));
}

double orc_limit(int limiter, double a, double b)
{
    return limiter == ORC_MINMOD ? minmod(a, b) : mc_lim(a, b);
}

/* ---- fast magnetosonic speed: muscl_hll.hpp:70-87 ---------------------------------------- */
double orc_fast_speed(const double* W, double gamma)
{
    const double inv_rho = 1.0 / W[R_RHO];
    const double cs2 = gamma * W[R_P] * inv_rho;
    const double bn2 = W[R_BN] * W[R_BN];
    const double b2 = bn2 + W[R_BT] * W[R_BT] + W[R_BZ] * W[R_BZ];
    const double va2 = b2 * inv_rho;
    const double van2 = bn2 * inv_rho;
    const double sum = cs2 + va2;
    const double disc = vmax(sum * sum - 4.0 * cs2 * van2, 0.0);
    const double cf2 = 0.5 * (sum + sqrt(disc));
    return sqrt(vmax(cf2, 0.0));
}

/* ---- conserved state + GLM flux along n: muscl_hll.hpp:89-124 --------------------------- */
static void cons_and_flux(const double* W, double inv_gm1, double ch2, double* U, double* F)
{
    const double mn = W[R_RHO] * W[R_VN];
    const double mt = W[R_RHO] * W[R_VT];
    const double mz = W[R_RHO] * W[R_VZ];
    const double v2 = W[R_VN] * W[R_VN] + W[R_VT] * W[R_VT] + W[R_VZ] * W[R_VZ];
    const double b2 = W[R_BN] * W[R_BN] + W[R_BT] * W[R_BT] + W[R_BZ] * W[R_BZ];
    const double pmag = 0.5 * b2;
    const double energy = W[R_P] * inv_gm1 + 0.5 * W[R_RHO] * v2 + pmag;
    const double pstar = W[R_P] + pmag;
    const double vb = W[R_VN] * W[R_BN] + W[R_VT] * W[R_BT] + W[R_VZ] * W[R_BZ];

    U[0] = W[R_RHO];
    U[1] = mn;
    U[2] = mt;
    U[3] = mz;
    U[4] = energy;
    U[5] = W[R_BN];
    U[6] = W[R_BT];
    U[7] = W[R_BZ];
    U[8] = W[R_PSI];

    F[0] = mn;
    F[1] = mn * W[R_VN] + pstar - W[R_BN] * W[R_BN];
    F[2] = mn * W[R_VT] - W[R_BN] * W[R_BT];
    F[3] = mn * W[R_VZ] - W[R_BN] * W[R_BZ];
    F[4] = (energy + pstar) * W[R_VN] - W[R_BN] * vb + W[R_PSI] * W[R_BN];
    F[5] = W[R_PSI];
    F[6] = W[R_VN] * W[R_BT] - W[R_VT] * W[R_BN];
    F[7] = W[R_VN] * W[R_BZ] - W[R_VZ] * W[R_BN];
    F[8] = ch2 * W[R_BN];
}

/* ---- MUSCL + HLL + GLM face flux: muscl_hll.hpp:59-68 (reconstruct), :126-209 ------------- */
void orc_face_flux(int limiter, const double* Wmm, const double* Wm, const double* W0,
                   const double* Wp, double gamma, double inv_gm1, double c_h, double* F)
{
    double WL[NC], WR[NC];
    for (int c = 0; c < NC; ++c) {
        const double dm = orc_limit(limiter, Wm[c] - Wmm[c], W0[c] - Wm[c]);
        const double d0 = orc_limit(limiter, W0[c] - Wm[c], Wp[c] - W0[c]);
        WL[c] = Wm[c] + 0.5 * dm;
        WR[c] = W0[c] - 0.5 * d0;
    }
    /* both sides fall back to first order unless rho, p of both are positive and finite */
    const int ok = (WL[R_RHO] > 0.0 && WL[R_P] > 0.0 && WL[R_RHO] < HUGE_VAL && WL[R_P] < HUGE_VAL) &&
                   (WR[R_RHO] > 0.0 && WR[R_P] > 0.0 && WR[R_RHO] < HUGE_VAL && WR[R_P] < HUGE_VAL);
    if (!ok) {
        for (int c = 0; c < NC; ++c) {
            WL[c] = Wm[c];
            WR[c] = W0[c];
        }
    }

    const double cfL = orc_fast_speed(WL, gamma);
    const double cfR = orc_fast_speed(WR, gamma);
    const double sL = vmin(vmin(WL[R_VN] - cfL, WR[R_VN] - cfR), -c_h);
    const double sR = vmax(vmax(WL[R_VN] + cfL, WR[R_VN] + cfR), c_h);

    const double ch2 = c_h * c_h;
    double UL[NC], UR[NC], FL[NC], FR[NC];
    cons_and_flux(WL, inv_gm1, ch2, UL, FL);
    cons_and_flux(WR, inv_gm1, ch2, UR, FR);

    const double denom = sR - sL;
    const double inv_denom = 1.0 / denom;
    const double sLsR = sL * sR;
    const int use_left = sL >= 0.0;
    const int use_right = sR <= 0.0;
    const int degenerate = fabs(denom) < 1.0e-14;
    for (int c = 0; c < NC; ++c) {
        const double hll = (sR * FL[c] - sL * FR[c] + sLsR * (UR[c] - UL[c])) * inv_denom;
        const double mid = degenerate ? 0.5 * (FL[c] + FR[c]) : hll;
        F[c] = use_left ? FL[c] : (use_right ? FR[c] : mid);
    }
}

/* ---- CFL signal speed of one cell: muscl_hll.hpp:211-239 ---------------------------------- */
double orc_cell_max_speed(const double* U9, double gamma)
{
    const double rho = U9[0], mx = U9[1], my = U9[2], mz = U9[3], e = U9[4];
    const double bx = U9[5], by = U9[6], bz = U9[7];
    const double inv_rho = 1.0 / rho;
    double W[NC];
    W[R_RHO] = rho;
    W[R_VN] = mx * inv_rho;
    W[R_VT] = my * inv_rho;
    W[R_VZ] = mz * inv_rho;
    const double ekin = 0.5 * (mx * mx + my * my + mz * mz) * inv_rho;
    const double emag = 0.5 * (bx * bx + by * by + bz * bz);
    W[R_P] = (gamma - 1.0) * (e - ekin - emag);
    W[R_BN] = bx;
    W[R_BT] = by;
    W[R_BZ] = bz;
    W[R_PSI] = 0.0;
    const double sx = fabs(W[R_VN]) + orc_fast_speed(W, gamma);
    double Wy[NC];
    memcpy(Wy, W, sizeof(W));
    Wy[R_VN] = W[R_VT];
    Wy[R_VT] = W[R_VN];
    Wy[R_BN] = by;
    Wy[R_BT] = bx;
    const double sy = fabs(Wy[R_VN]) + orc_fast_speed(Wy, gamma);
    return vmax(sx, sy);
}

/* ---- non-ideal electric field: include/mhd/nonideal.hpp:63-102 ----------------------------- */
void orc_nonideal_e(const double* J, const double* B, const orc_params* p, double* E)
{
    E[0] = 0.0;
    E[1] = 0.0;
    E[2] = 0.0;
    if (p->ohmic) {
        E[0] += p->eta_ohm * J[0];
        E[1] += p->eta_ohm * J[1];
        E[2] += p->eta_ohm * J[2];
    }
    if (!(p->hall || p->ambipolar)) {
        return;
    }
    const double B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    if (B2 < 1.0e-14) {
        return;
    }
    /* J x B */
    const double jb0 = J[1] * B[2] - J[2] * B[1];
    const double jb1 = J[2] * B[0] - J[0] * B[2];
    const double jb2 = J[0] * B[1] - J[1] * B[0];
    if (p->hall) {
        const double invB = 1.0 / sqrt(B2);
        E[0] += p->eta_hall * jb0 * invB;
        E[1] += p->eta_hall * jb1 * invB;
        E[2] += p->eta_hall * jb2 * invB;
    }
    if (p->ambipolar) {
        /* (J x B) x B */
        const double a0 = jb1 * B[2] - jb2 * B[1];
        const double a1 = jb2 * B[0] - jb0 * B[2];
        const double a2 = jb0 * B[1] - jb1 * B[0];
        const double invB2 = 1.0 / B2;
        E[0] += p->eta_ambipolar * a0 * invB2;
        E[1] += p->eta_ambipolar * a1 * invB2;
        E[2] += p->eta_ambipolar * a2 * invB2;
    }
}

static int nonideal_any(const orc_params* p)
{
    /* include/physics_config/mhd_config.hpp:90-94 */
    return (p->ohmic && p->eta_ohm != 0.0) || (p->hall && p->eta_hall != 0.0) ||
           (p->ambipolar && p->eta_ambipolar != 0.0);
}

static double max_diffusivity(const orc_params* p)
{
    /* include/physics_config/mhd_config.hpp:97-110 (std::max(a, b) = a < b ? b : a) */
    double eta = 0.0;
    if (p->ohmic) {
        eta = vmax(eta, fabs(p->eta_ohm));
    }
    if (p->hall) {
        eta = vmax(eta, fabs(p->eta_hall));
    }
    if (p->ambipolar) {
        eta = vmax(eta, fabs(p->eta_ambipolar));
    }
    return eta;
}

/* ---- grid helpers: include/grid/grid.hpp:24-25,47-71 --------------------------------------- */
typedef struct {
    int nx, ny, ng, nxt, nyt;
    size_t n;
    double dx, dy;
} geom;

static geom make_geom(const orc_grid* g)
{
    geom q;
    q.nx = g->nx;
    q.ny = g->ny;
    q.ng = g->ng;
    q.nxt = g->nx + 2 * g->ng;
    q.nyt = g->ny + 2 * g->ng;
    q.n = (size_t)q.nxt * (size_t)q.nyt;
    q.dx = (g->x_max - g->x_min) / (double)g->nx;
    q.dy = (g->y_max - g->y_min) / (double)g->ny;
    return q;
}

/* ---- boundary conditions: include/boundary_conditions/boundary_conditions.hpp:26-103 ------- */
int orc_periodic_source(int g, int n, int ng, int low)
{
    const int offset = low ? (g - ng) : g;
    return ng + ((offset % n) + n) % n;
}

static void copy_cell(double* U, size_t n, size_t dst, size_t src)
{
    for (int c = 0; c < NC; ++c) {
        U[c * n + dst] = U[c * n + src];
    }
}

int orc_apply_bc(const orc_grid* g, double* U, int bc_x, int bc_y)
{
    const geom q = make_geom(g);
    if (q.ng < 1) {
        return ORC_EINVAL;
    }
    if (bc_x == ORC_REFLECTIVE || bc_y == ORC_REFLECTIVE) {
        return ORC_EINVAL;
    }
    /* x pass over every row (ghost rows included) */
    for (int j = 0; j < q.nyt; ++j) {
        const size_t row = (size_t)j * q.nxt;
        for (int gh = 0; gh < q.ng; ++gh) {
            const int il = gh;
            const int ir = q.ng + q.nx + gh;
            int sl, sr;
            if (bc_x == ORC_PERIODIC) {
                sl = orc_periodic_source(gh, q.nx, q.ng, 1);
                sr = orc_periodic_source(gh, q.nx, q.ng, 0);
            } else {
                sl = q.ng;
                sr = q.ng + q.nx - 1;
            }
            copy_cell(U, q.n, row + il, row + sl);
            copy_cell(U, q.n, row + ir, row + sr);
        }
    }
    /* y pass over the full x range (fills corners) */
    for (int i = 0; i < q.nxt; ++i) {
        for (int gh = 0; gh < q.ng; ++gh) {
            const int jb = gh;
            const int jt = q.ng + q.ny + gh;
            int sb, st;
            if (bc_y == ORC_PERIODIC) {
                sb = orc_periodic_source(gh, q.ny, q.ng, 1);
                st = orc_periodic_source(gh, q.ny, q.ng, 0);
            } else {
                sb = q.ng;
                st = q.ng + q.ny - 1;
            }
            copy_cell(U, q.n, (size_t)jb * q.nxt + i, (size_t)sb * q.nxt + i);
            copy_cell(U, q.n, (size_t)jt * q.nxt + i, (size_t)st * q.nxt + i);
        }
    }
    return ORC_OK;
}

/* ---- scalar conversions used by the wall fill: include/mhd/mhd_physics.hpp:23-59 ---------- */
static void to_prim_scalar(const double* S, double gamma, double* W)
{
    W[0] = S[0];
    W[1] = S[1] / S[0];
    W[2] = S[2] / S[0];
    W[3] = S[3] / S[0];
    W[5] = S[5];
    W[6] = S[6];
    W[7] = S[7];
    W[8] = S[8];
    const double e_kin = 0.5 * (S[1] * S[1] + S[2] * S[2] + S[3] * S[3]) / S[0];
    const double e_mag = 0.5 * (S[5] * S[5] + S[6] * S[6] + S[7] * S[7]);
    W[4] = (gamma - 1.0) * (S[4] - e_kin - e_mag);
}

static void to_cons_scalar(const double* W, double gamma, double* S)
{
    S[0] = W[0];
    S[1] = W[0] * W[1];
    S[2] = W[0] * W[2];
    S[3] = W[0] * W[3];
    S[5] = W[5];
    S[6] = W[6];
    S[7] = W[7];
    S[8] = W[8];
    const double e_kin = 0.5 * W[0] * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3]);
    const double e_mag = 0.5 * (W[5] * W[5] + W[6] * W[6] + W[7] * W[7]);
    S[4] = W[4] / (gamma - 1.0) + e_kin + e_mag;
}

/* ---- wall fill: include/mhd/mask.hpp:85-149 ------------------------------------------------ */
void orc_wall_fill(const orc_grid* g, double* U, const uint8_t* mask, double gamma)
{
    const geom q = make_geom(g);
    if (mask == NULL) {
        return;
    }
    const size_t ncell = (size_t)q.nx * (size_t)q.ny;
    int all_fluid = 1;
    for (size_t k = 0; k < ncell; ++k) {
        if (mask[k] == 0) {
            all_fluid = 0;
            break;
        }
    }
    if (all_fluid) {
        return;
    }
    for (int j = q.ng; j < q.ng + q.ny; ++j) {
        for (int i = q.ng; i < q.ng + q.nx; ++i) {
            if (mask[(size_t)(j - q.ng) * q.nx + (i - q.ng)] != 0) {
                continue;
            }
            const int cand[4][3] = {{i + 1, j, 0}, {i - 1, j, 0}, {i, j + 1, 1}, {i, j - 1, 1}};
            int filled = 0;
            for (int k = 0; k < 4 && !filled; ++k) {
                const int ni = cand[k][0], nj = cand[k][1], axis = cand[k][2];
                if (ni < q.ng || ni >= q.ng + q.nx || nj < q.ng || nj >= q.ng + q.ny) {
                    continue;
                }
                if (mask[(size_t)(nj - q.ng) * q.nx + (ni - q.ng)] == 0) {
                    continue;
                }
                const size_t src = (size_t)nj * q.nxt + ni;
                double S[NC], W[NC], O[NC];
                for (int c = 0; c < NC; ++c) {
                    S[c] = U[c * q.n + src];
                }
                to_prim_scalar(S, gamma, W);
                if (axis == 0) {
                    W[1] = -W[1];
                } else {
                    W[2] = -W[2];
                }
                if (W[0] < 1e-12) {
                    W[0] = 1e-12;
                }
                if (W[4] < 1e-12) {
                    W[4] = 1e-12;
                }
                to_cons_scalar(W, gamma, O);
                const size_t dst = (size_t)j * q.nxt + i;
                for (int c = 0; c < NC; ++c) {
                    U[c * q.n + dst] = O[c];
                }
                filled = 1;
            }
            if (!filled) {
                double W[NC] = {1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
                double O[NC];
                to_cons_scalar(W, gamma, O);
                const size_t dst = (size_t)j * q.nxt + i;
                for (int c = 0; c < NC; ++c) {
                    U[c * q.n + dst] = O[c];
                }
            }
        }
    }
}

static int mask_active(const orc_grid* g, const uint8_t* mask)
{
    if (mask == NULL) {
        return 0;
    }
    const size_t ncell = (size_t)g->nx * (size_t)g->ny;
    for (size_t k = 0; k < ncell; ++k) {
        if (mask[k] == 0) {
            return 1;
        }
    }
    return 0;
}

/* ---- primitive recovery: include/mhd/kernels/cpu_rows.hpp:130-155 --------------------------- */
static void fill_primitives(const geom* q, const double* U, double* P, double gamma)
{
    const size_t n = q->n;
    const double gm1 = gamma - 1.0;
#pragma omp parallel for schedule(static)
    for (long long kk = 0; kk < (long long)n; ++kk) {
        const size_t k = (size_t)kk;
        const double rho = U[0 * n + k];
        const double mx = U[1 * n + k];
        const double my = U[2 * n + k];
        const double mz = U[3 * n + k];
        const double e = U[4 * n + k];
        const double bx = U[5 * n + k];
        const double by = U[6 * n + k];
        const double bz = U[7 * n + k];
        const double inv_rho = 1.0 / rho;
        const double ekin = 0.5 * (mx * mx + my * my + mz * mz) * inv_rho;
        const double emag = 0.5 * (bx * bx + by * by + bz * bz);
        P[0 * n + k] = rho;
        P[1 * n + k] = mx * inv_rho;
        P[2 * n + k] = my * inv_rho;
        P[3 * n + k] = mz * inv_rho;
        P[4 * n + k] = gm1 * (e - ekin - emag);
        P[5 * n + k] = bx;
        P[6 * n + k] = by;
        P[7 * n + k] = bz;
        P[8 * n + k] = U[8 * n + k];
    }
}

/* rotated load: include/mhd/kernels/cpu_rows.hpp:40-61 */
static inline void load_rot(const double* P, size_t n, size_t k, int ydir, double* W)
{
    W[R_RHO] = P[0 * n + k];
    W[R_VZ] = P[3 * n + k];
    W[R_P] = P[4 * n + k];
    W[R_BZ] = P[7 * n + k];
    W[R_PSI] = P[8 * n + k];
    if (!ydir) {
        W[R_VN] = P[1 * n + k];
        W[R_VT] = P[2 * n + k];
        W[R_BN] = P[5 * n + k];
        W[R_BT] = P[6 * n + k];
    } else {
        W[R_VN] = P[2 * n + k];
        W[R_VT] = P[1 * n + k];
        W[R_BN] = P[6 * n + k];
        W[R_BT] = P[5 * n + k];
    }
}

/* un-rotated store: include/mhd/kernels/cpu_rows.hpp:63-82 */
static inline void unrotate(const double* Fr, int ydir, double* F)
{
    F[C_RHO] = Fr[0];
    F[C_MZ] = Fr[3];
    F[C_E] = Fr[4];
    F[C_BZ] = Fr[7];
    F[C_PSI] = Fr[8];
    if (!ydir) {
        F[C_MX] = Fr[1];
        F[C_MY] = Fr[2];
        F[C_BX] = Fr[5];
        F[C_BY] = Fr[6];
    } else {
        F[C_MX] = Fr[2];
        F[C_MY] = Fr[1];
        F[C_BX] = Fr[6];
        F[C_BY] = Fr[5];
    }
}

/* ---- current J = curl B: include/mhd/nonideal.hpp:30-60 ------------------------------------ */
static void compute_current(const geom* q, const double* P, double* J)
{
    const size_t n = q->n;
    const double inv_2dx = 0.5 / q->dx;
    const double inv_2dy = 0.5 / q->dy;
    const double* bx = P + 5 * n;
    const double* by = P + 6 * n;
    const double* bz = P + 7 * n;
    const int nxt = q->nxt;
#pragma omp parallel for schedule(static)
    for (int j = 1; j < q->nyt - 1; ++j) {
        for (int i = 1; i < nxt - 1; ++i) {
            const size_t k = (size_t)j * nxt + i;
            const double dbz_dy = (bz[k + nxt] - bz[k - nxt]) * inv_2dy;
            const double dbz_dx = (bz[k + 1] - bz[k - 1]) * inv_2dx;
            const double dby_dx = (by[k + 1] - by[k - 1]) * inv_2dx;
            const double dbx_dy = (bx[k + nxt] - bx[k - nxt]) * inv_2dy;
            J[0 * n + k] = dbz_dy;
            J[1 * n + k] = -dbz_dx;
            J[2 * n + k] = dby_dx - dbx_dy;
        }
    }
}

static inline void face_avg(const double* A, size_t n, size_t a, size_t b, double* out)
{
    /* avg3: include/mhd/nonideal.hpp:104-107 */
    out[0] = 0.5 * (A[0 * n + a] + A[0 * n + b]);
    out[1] = 0.5 * (A[1 * n + a] + A[1 * n + b]);
    out[2] = 0.5 * (A[2 * n + a] + A[2 * n + b]);
}

/* ---- spatial operator: include/mhd/rhs.hpp:116-123 (assembly), :140-240 -------------------- */
typedef struct {
    double* P;  /* 9n primitives */
    double* J;  /* 3n current */
    double* FX; /* 9 * ny * (nx+1) x-face fluxes, face (i - 1/2) of interior row */
    double* FY; /* 9 * (ny+1) * nx y-face fluxes */
} rhs_work;

static void rhs_with_work(const orc_grid* g, const double* U, double* rhs, const orc_params* p,
                          double c_h, const uint8_t* mask, rhs_work* w)
{
    const geom q = make_geom(g);
    const size_t n = q.n;
    const double gamma = p->gamma;
    const double inv_gm1 = 1.0 / (gamma - 1.0);
    const int ni = nonideal_any(p);
    const int lim = p->limiter;
    fill_primitives(&q, U, w->P, gamma);
    if (ni) {
        compute_current(&q, w->P, w->J);
    }
    const int fxw = q.nx + 1;
    const size_t nfx = (size_t)q.ny * fxw;
    const size_t nfy = (size_t)(q.ny + 1) * q.nx;
    const double* Bp = w->P + 5 * n; /* bx, by, bz planes of the primitive field */

#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < q.ny; ++jj) {
        const int j = q.ng + jj;
        for (int k = 0; k <= q.nx; ++k) {
            const int i = q.ng + k;
            const size_t c0 = (size_t)j * q.nxt + i;
            double Wmm[NC], Wm[NC], W0[NC], Wp[NC], Fr[NC], F[NC];
            load_rot(w->P, n, c0 - 2, 0, Wmm);
            load_rot(w->P, n, c0 - 1, 0, Wm);
            load_rot(w->P, n, c0, 0, W0);
            load_rot(w->P, n, c0 + 1, 0, Wp);
            orc_face_flux(lim, Wmm, Wm, W0, Wp, gamma, inv_gm1, c_h, Fr);
            unrotate(Fr, 0, F);
            if (ni) {
                /* include/mhd/rhs.hpp:74-93 */
                double Jf[3], Bf[3], E[3];
                face_avg(w->J, n, c0 - 1, c0, Jf);
                face_avg(Bp, n, c0 - 1, c0, Bf);
                orc_nonideal_e(Jf, Bf, p, E);
                F[C_BY] += -E[2];
                F[C_BZ] += E[1];
                F[C_E] += Bf[1] * E[2] - Bf[2] * E[1];
            }
            for (int c = 0; c < NC; ++c) {
                w->FX[c * nfx + (size_t)jj * fxw + k] = F[c];
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj <= q.ny; ++jj) {
        const int j = q.ng + jj;
        for (int k = 0; k < q.nx; ++k) {
            const int i = q.ng + k;
            const size_t c0 = (size_t)j * q.nxt + i;
            const size_t s = (size_t)q.nxt;
            double Wmm[NC], Wm[NC], W0[NC], Wp[NC], Fr[NC], F[NC];
            load_rot(w->P, n, c0 - 2 * s, 1, Wmm);
            load_rot(w->P, n, c0 - s, 1, Wm);
            load_rot(w->P, n, c0, 1, W0);
            load_rot(w->P, n, c0 + s, 1, Wp);
            orc_face_flux(lim, Wmm, Wm, W0, Wp, gamma, inv_gm1, c_h, Fr);
            unrotate(Fr, 1, F);
            if (ni) {
                /* include/mhd/rhs.hpp:96-114 */
                double Jf[3], Bf[3], E[3];
                face_avg(w->J, n, c0 - s, c0, Jf);
                face_avg(Bp, n, c0 - s, c0, Bf);
                orc_nonideal_e(Jf, Bf, p, E);
                F[C_BX] += E[2];
                F[C_BZ] += -E[0];
                F[C_E] += Bf[2] * E[0] - Bf[0] * E[2];
            }
            for (int c = 0; c < NC; ++c) {
                w->FY[c * nfy + (size_t)jj * q.nx + k] = F[c];
            }
        }
    }

    const double inv_dx = 1.0 / q.dx;
    const double inv_dy = 1.0 / q.dy;
    const double h = vmin(q.dx, q.dy);
    const double psi_damp = (c_h > 0.0 && p->glm_alpha > 0.0) ? (c_h * p->glm_alpha / h) : 0.0;
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < q.ny; ++jj) {
        const int j = q.ng + jj;
        for (int k = 0; k < q.nx; ++k) {
            const size_t cell = (size_t)j * q.nxt + q.ng + k;
            for (int c = 0; c < NC; ++c) {
                const double* fx = w->FX + c * nfx + (size_t)jj * fxw;
                const double* gb = w->FY + c * nfy + (size_t)jj * q.nx;
                const double* gt = gb + q.nx;
                rhs[c * n + cell] = -(fx[k + 1] - fx[k]) * inv_dx - (gt[k] - gb[k]) * inv_dy;
            }
            if (psi_damp != 0.0) {
                rhs[8 * n + cell] -= psi_damp * U[8 * n + cell];
            }
            if (mask != NULL && mask[(size_t)jj * q.nx + k] == 0) {
                for (int c = 0; c < NC; ++c) {
                    rhs[c * n + cell] = 0.0;
                }
            }
        }
    }
}

static int alloc_work(const geom* q, rhs_work* w)
{
    const size_t n = q->n;
    w->P = (double*)calloc(NC * n, sizeof(double));
    w->J = (double*)calloc(3 * n, sizeof(double));
    w->FX = (double*)calloc((size_t)NC * q->ny * (q->nx + 1), sizeof(double));
    w->FY = (double*)calloc((size_t)NC * (q->ny + 1) * q->nx, sizeof(double));
    return (w->P && w->J && w->FX && w->FY) ? 0 : -1;
}

static void free_work(rhs_work* w)
{
    free(w->P);
    free(w->J);
    free(w->FX);
    free(w->FY);
}

void orc_rhs(const orc_grid* g, const double* U, double* rhs, const orc_params* p, double c_h,
             const uint8_t* mask)
{
    const geom q = make_geom(g);
    rhs_work w;
    if (alloc_work(&q, &w) != 0) {
        free_work(&w);
        return;
    }
    rhs_with_work(g, U, rhs, p, c_h, mask_active(g, mask) ? mask : NULL, &w);
    free_work(&w);
}

/* ---- CFL: include/mhd/cfl.hpp:28-92 ------------------------------------------------------- */
/* include/initial_condition/inititial_condition.hpp:273-287 */
double orc_max_div_b(const orc_grid* g, const double* U)
{
    const geom q = make_geom(g);
    const double inv_2dx = 0.5 / q.dx;
    const double inv_2dy = 0.5 / q.dy;
    const double* bx = U + 5 * q.n;
    const double* by = U + 6 * q.n;
    double max_abs = 0.0;
    for (int j = q.ng; j < q.ng + q.ny; ++j) {
        for (int i = q.ng; i < q.ng + q.nx; ++i) {
            const size_t c = (size_t)j * q.nxt + i;
            const double dbx = (bx[c + 1] - bx[c - 1]) * inv_2dx;
            const double dby = (by[c + q.nxt] - by[c - q.nxt]) * inv_2dy;
            max_abs = vmax(max_abs, fabs(dbx + dby));
        }
    }
    return max_abs;
}

double orc_cfl(const orc_grid* g, const double* U, const orc_params* p, const uint8_t* mask,
               double* c_h_out)
{
    const geom q = make_geom(g);
    const size_t n = q.n;
    const double inv_dx = 1.0 / q.dx;
    const double inv_dy = 1.0 / q.dy;
    const uint8_t* m = mask_active(g, mask) ? mask : NULL;
    double c_h = 0.0;
#pragma omp parallel
    {
        double local = 0.0;
#pragma omp for schedule(static)
        for (int jj = 0; jj < q.ny; ++jj) {
            const int j = q.ng + jj;
            for (int k = 0; k < q.nx; ++k) {
                if (m != NULL && m[(size_t)jj * q.nx + k] == 0) {
                    continue;
                }
                const size_t cell = (size_t)j * q.nxt + q.ng + k;
                double U9[NC];
                for (int c = 0; c < NC; ++c) {
                    U9[c] = U[c * n + cell];
                }
                local = vmax(local, orc_cell_max_speed(U9, p->gamma));
            }
        }
#pragma omp critical
        {
            if (local > c_h) {
                c_h = local;
            }
        }
    }
    *c_h_out = c_h;
    double dt;
    if (c_h <= 0.0) {
        dt = HUGE_VAL;
    } else {
        dt = p->cfl / (c_h * inv_dx + c_h * inv_dy);
    }
    const double eta_max = max_diffusivity(p);
    if (eta_max > 0.0) {
        const double h = vmin(q.dx, q.dy);
        const double dt_diff = p->cfl * h * h / (2.0 * 2.0 * eta_max);
        dt = vmin(dt, dt_diff);
    }
    return dt;
}

/* ---- SSP-RK2 and the solve loop: include/mhd/time_integrator.hpp:36-106,
 *      include/mhd/field_ops.hpp:47-63 (x + a*y, 0.5*(u+s) + (0.5*dt)*r) ---------------- */
/* The solve loop; with `forced` it takes exactly `nforced` steps whose dt are forced[k]
 * instead of the CFL dt (c_h of each step still comes from the CFL reduction). */
static int solve_impl(const orc_grid* g, int bc_x, int bc_y, const uint8_t* mask,
                      int reflective_walls, const orc_params* p, int allow_max_steps, double* U,
                      double* dt_log, int dt_cap, const double* forced, int nforced,
                      double* t_out, int* steps_out, double* ch_out)
{
    if (p->t_end < 0.0 || p->cfl <= 0.0 || g->ng < 2 || g->nx < 1 || g->ny < 1) {
        return ORC_EINVAL;
    }
    if (bc_x == ORC_REFLECTIVE || bc_y == ORC_REFLECTIVE) {
        return ORC_EINVAL;
    }
    const geom q = make_geom(g);
    const size_t n = q.n;
    const uint8_t* m = mask_active(g, mask) ? mask : NULL;
    const int fill = (m != NULL && reflective_walls);

    double* Us = (double*)calloc(NC * n, sizeof(double));
    double* R = (double*)calloc(NC * n, sizeof(double));
    rhs_work w;
    if (Us == NULL || R == NULL || alloc_work(&q, &w) != 0) {
        free(Us);
        free(R);
        return ORC_EINVAL;
    }

    double t = 0.0;
    int steps = 0;
    double c_h_last = 0.0;
    while (forced != NULL ? steps < nforced : (t < p->t_end && steps < p->max_steps)) {
        double c_h;
        const double dt_cfl = orc_cfl(g, U, p, m, &c_h);
        c_h_last = c_h;
        const double rem = p->t_end - t;
        const double dt = forced != NULL ? forced[steps] : vmin(dt_cfl, rem);
        if (dt_log != NULL && steps < dt_cap) {
            dt_log[steps] = dt;
        }
        /* stage 1 */
        orc_apply_bc(g, U, bc_x, bc_y);
        if (fill) {
            orc_wall_fill(g, U, m, p->gamma);
        }
        rhs_with_work(g, U, R, p, c_h, m, &w);
#pragma omp parallel for schedule(static)
        for (long long kk = 0; kk < (long long)(NC * n); ++kk) {
            Us[kk] = U[kk] + dt * R[kk];
        }
        /* stage 2 */
        orc_apply_bc(g, Us, bc_x, bc_y);
        if (fill) {
            orc_wall_fill(g, Us, m, p->gamma);
        }
        rhs_with_work(g, Us, R, p, c_h, m, &w);
        const double half_dt = 0.5 * dt;
#pragma omp parallel for schedule(static)
        for (long long kk = 0; kk < (long long)(NC * n); ++kk) {
            U[kk] = 0.5 * (U[kk] + Us[kk]) + half_dt * R[kk];
        }
        t += dt;
        ++steps;
    }
    int rc = ORC_OK;
    if (forced == NULL && t < p->t_end && steps >= p->max_steps && !allow_max_steps) {
        rc = ORC_EMAXSTEPS;
    }
    orc_apply_bc(g, U, bc_x, bc_y);
    if (fill) {
        orc_wall_fill(g, U, m, p->gamma);
    }
    free(Us);
    free(R);
    free_work(&w);
    if (t_out) {
        *t_out = t;
    }
    if (steps_out) {
        *steps_out = steps;
    }
    if (ch_out) {
        *ch_out = c_h_last;
    }
    return rc;
}

/* include/mhd/time_integrator.hpp:62-106 (solve) */
int orc_solve(const orc_grid* g, int bc_x, int bc_y, const uint8_t* mask, int reflective_walls,
              const orc_params* p, int allow_max_steps, double* U, double* dt_log, int dt_cap,
              double* t_out, int* steps_out, double* ch_out)
{
    return solve_impl(g, bc_x, bc_y, mask, reflective_walls, p, allow_max_steps, U, dt_log,
                      dt_cap, NULL, 0, t_out, steps_out, ch_out);
}

/* The same steps with a given dt sequence (time_integrator.hpp:36-60 per step): compares a
 * solver whose dt may differ in the last bits (a contracted fp64 build) at its own dt. */
int orc_solve_dt(const orc_grid* g, int bc_x, int bc_y, const uint8_t* mask, int reflective_walls,
                 const orc_params* p, double* U, const double* dts, int nsteps, double* t_out,
                 int* steps_out, double* ch_out)
{
    if (dts == NULL || nsteps < 0) {
        return ORC_EINVAL;
    }
    return solve_impl(g, bc_x, bc_y, mask, reflective_walls, p, 1, U, NULL, 0, dts, nsteps,
                      t_out, steps_out, ch_out);
}

/* ---- reference-named initial conditions: include/initial_condition/inititial_condition.hpp
 *      :28-192 (IC formulas), :194-224 (names), :226-270 (fill); MHDSolver::initialize
 *      include/mhd/solver_api.hpp:276-292 (Sod gamma 1.4, Brio-Wu gamma 2 are set by the
 *      caller before this fill) -------------------------------------------------------------- */
static void ic_point(int kind, double x, double y, double* W)
{
    for (int c = 0; c < NC; ++c) {
        W[c] = 0.0;
    }
    switch (kind) {
    case 0: /* uniform */
        W[0] = 1.0;
        W[4] = 1.0;
        break;
    case 1: /* sine */
        W[0] = 1.0 + 0.1 * sin(2.0 * kPi * x);
        W[4] = 1.0;
        break;
    case 2: { /* blast */
        const double r = hypot(x - 0.5, y - 0.5);
        W[0] = 1.0;
        W[4] = (r < 0.1) ? 10.0 : 0.1;
        W[5] = 1.0 / sqrt(2.0);
        W[6] = 1.0 / sqrt(2.0);
        break;
    }
    case 3: /* Orszag-Tang, B0 = 1 */
        W[0] = 25.0 / (36.0 * kPi);
        W[1] = -sin(2.0 * kPi * y);
        W[2] = sin(2.0 * kPi * x);
        W[4] = 5.0 / (12.0 * kPi);
        W[5] = -sin(2.0 * kPi * y);
        W[6] = sin(4.0 * kPi * x);
        break;
    case 4: /* Sod */
        if (x < 0.5) {
            W[0] = 1.0;
            W[4] = 1.0;
        } else {
            W[0] = 0.125;
            W[4] = 0.1;
        }
        break;
    case 5: /* Brio-Wu */
        W[5] = 0.75;
        if (x < 0.5) {
            W[0] = 1.0;
            W[4] = 1.0;
            W[6] = 1.0;
        } else {
            W[0] = 0.125;
            W[4] = 0.1;
            W[6] = -1.0;
        }
        break;
    case 6: { /* rotor */
        const double ddx = x - 0.5;
        const double ddy = y - 0.5;
        const double r = hypot(ddx, ddy);
        W[5] = 5.0 / sqrt(4.0 * kPi);
        W[4] = 1.0;
        double f;
        if (r <= 0.1) {
            f = 1.0;
            W[0] = 10.0;
        } else if (r >= 0.115) {
            f = 0.0;
            W[0] = 1.0;
        } else {
            f = (0.115 - r) / (0.115 - 0.1);
            W[0] = 1.0 + 9.0 * f;
        }
        if (r > 0.0) {
            W[1] = -f * 2.0 * ddy / 0.1;
            W[2] = f * 2.0 * ddx / 0.1;
        }
        break;
    }
    case 7: { /* Kelvin-Helmholtz double shear */
        const int in = (y > 0.25 && y < 0.75);
        W[0] = in ? 2.0 : 1.0;
        W[1] = in ? 0.5 : -0.5;
        W[2] = 0.01 * sin(2.0 * kPi * x);
        W[4] = 2.5;
        break;
    }
    case 8: { /* circularly polarized Alfven wave, delta = 1e-3 */
        const double delta = 1.0e-3;
        const double phase = 2.0 * kPi * x;
        W[0] = 1.0;
        W[4] = 0.1;
        W[5] = 1.0;
        W[6] = delta * cos(phase);
        W[7] = delta * sin(phase);
        W[2] = -(delta / sqrt(1.0)) * cos(phase);
        W[3] = -(delta / sqrt(1.0)) * sin(phase);
        break;
    }
    default:
        break;
    }
}

static int ic_kind(const char* s)
{
    if (!strcmp(s, "uniform") || !strcmp(s, "Uniform")) return 0;
    if (!strcmp(s, "sine") || !strcmp(s, "SineWave")) return 1;
    if (!strcmp(s, "blast") || !strcmp(s, "BlastWave")) return 2;
    if (!strcmp(s, "orszag_tang") || !strcmp(s, "OrszagTang")) return 3;
    if (!strcmp(s, "sod") || !strcmp(s, "Sod")) return 4;
    if (!strcmp(s, "brio_wu") || !strcmp(s, "BrioWu")) return 5;
    if (!strcmp(s, "rotor") || !strcmp(s, "Rotor")) return 6;
    if (!strcmp(s, "kelvin_helmholtz") || !strcmp(s, "KelvinHelmholtz") || !strcmp(s, "kh")) return 7;
    if (!strcmp(s, "alfven") || !strcmp(s, "alfven_wave") || !strcmp(s, "AlfvenWave")) return 8;
    return -1;
}

int orc_init_ic(const char* name, const orc_grid* g, double gamma, double* U)
{
    const int kind = ic_kind(name);
    if (kind < 0) {
        return ORC_EINVAL;
    }
    const geom q = make_geom(g);
    for (int j = q.ng; j < q.ng + q.ny; ++j) {
        for (int i = q.ng; i < q.ng + q.nx; ++i) {
            const double x = g->x_min + (i - q.ng + 0.5) * q.dx;
            const double y = g->y_min + (j - q.ng + 0.5) * q.dy;
            double W[NC], S[NC];
            ic_point(kind, x, y, W);
            to_cons_scalar(W, gamma, S);
            const size_t k = (size_t)j * q.nxt + i;
            for (int c = 0; c < NC; ++c) {
                U[c * q.n + k] = S[c];
            }
        }
    }
    return ORC_OK;
}
