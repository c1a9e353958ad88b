// Device point physics of the GLM-MHD MUSCL + HLL scheme, written for CDNA (gfx950).
//
// Every routine keeps the reference CPU kernels' operation order (include/mhd/kernels/
// muscl_hll.hpp:33-239, include/mhd/kernels/cpu_rows.hpp:130-155, include/mhd/nonideal.hpp:63-102)
// so that the fp64 instantiation, compiled with -ffp-contract=off, rounds exactly like the CPU path.
//
// Values are kept in the physical component order (rho, vx, vy, vz, p, bx, by, bz, psi).
// A face along direction DIR (0 = x, 1 = y) reads its normal/tangential slots through the
// compile-time indices of FaceDir<DIR>, which replaces the reference's load/store rotation.
//
// Precision policy (template parameter R):
//   double: everything in fp64, in the reference's operation order.
//   float : bulk arithmetic in fp32. Pressure recovery from total energy and the fast-speed
//           discriminant are evaluated in fp64 (MixedOps<float>), because both subtract nearly
//           equal quantities (low plasma beta or high Mach number; strong field along the
//           normal); the HLL flux uses clamped wave speeds (same values up to rounding, fewer
//           instructions).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mhdhip {

constexpr int kNC = 9;
enum Comp { RHO = 0, VX = 1, VY = 2, VZ = 3, PR = 4, BX = 5, BY = 6, BZ = 7, PSI = 8 };
enum Limiter { kMinmod = 0, kMC = 1 };

template <int DIR>
struct FaceDir {
    static constexpr int N = DIR ? VY : VX;   // normal velocity / momentum slot
    static constexpr int T = DIR ? VX : VY;   // tangential velocity / momentum slot
    static constexpr int BN = DIR ? BY : BX;  // normal field slot
    static constexpr int BT = DIR ? BX : BY;  // tangential field slot
};

template <typename R>
struct Lit;
template <>
struct Lit<double> {
    static constexpr double kDegenerate = 1.0e-14;
};
template <>
struct Lit<float> {
    static constexpr float kDegenerate = 1.0e-14f;
};

// min/max with the reference's tie and NaN rule (include/mhd/simd.hpp:40-47)
template <typename R>
__device__ __forceinline__ R vmin(R a, R b) { return (b < a) ? b : a; }
template <typename R>
__device__ __forceinline__ R vmax(R a, R b) { return (a < b) ? b : a; }
// min/max of the flux's wave speeds and discriminants: the select rule above in the exact fp64
// build; in the fp32 and contracted fp64 builds the hardware v_min / v_max (fused into v_min3 /
// v_max3 where they chain), which differ from it only in which zero or non-NaN operand they
// return. Step control (dt, the CFL reductions) keeps the select rule in every build.
template <typename R>
__device__ __forceinline__ R hmin(R a, R b) { return vmin(a, b); }
template <typename R>
__device__ __forceinline__ R hmax(R a, R b) { return vmax(a, b); }
template <>
__device__ __forceinline__ float hmin(float a, float b) { return __builtin_fminf(a, b); }
template <>
__device__ __forceinline__ float hmax(float a, float b) { return __builtin_fmaxf(a, b); }

// fp64: correctly rounded sqrt and division (parity with the CPU path).
// fp64, fast build (MHD_FAST_FP64, its own translation unit, compiled with contraction): the
// hardware reciprocal and reciprocal square root (v_rcp_f64, v_rsq_f64) refined by Newton /
// Goldschmidt iterations to about 1 ulp, held to 1e-12 relative L1 against the exact path.
// fp32-bulk: the hardware square root and reciprocal (v_sqrt_f32, v_rcp_f32, ~1 ulp); the FP32
// path is held to a relative L1 tolerance against fp64, not to bit parity.
#ifndef MHD_FAST_FP64
#define MHD_FAST_FP64 0
#endif
#if MHD_FAST_FP64
template <>
__device__ __forceinline__ double hmin(double a, double b) { return __builtin_fmin(a, b); }
template <>
__device__ __forceinline__ double hmax(double a, double b) { return __builtin_fmax(a, b); }

__device__ __forceinline__ double dsqrt(double x)
{
    // y ~ 1/sqrt(x); one coupled iteration on g ~ sqrt(x), h ~ 1/(2 sqrt(x)), then the
    // residual x - g^2 corrects g
    const double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * y;
    double g = x * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    // 0, +inf and NaN pass through; the arguments here are never negative
    return (x > 0.0 && x < __builtin_huge_val()) ? g : x;
}
__device__ __forceinline__ double drcp(double x)
{
    // two Newton steps r <- r + r (1 - x r) from the hardware estimate
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
#else
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ double drcp(double x) { return 1.0 / x; }
#endif
__device__ __forceinline__ float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float drcp(float x) { return __builtin_amdgcn_rcpf(x); }
// a b + c: one fused multiply-add in the FP32 and contracted fp64 builds, the product and the
// sum rounded separately in the exact fp64 build. The contracted builds fuse only within one
// expression (-ffp-contract=on); a product and a sum in separate statements are joined here.
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c)
{
    if constexpr (MHD_FAST_FP64) {
        return __builtin_fma(a, b, c);
    } else {
        return a * b + c;
    }
}
__device__ __forceinline__ double dabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float dabs(float x) { return __builtin_fabsf(x); }

// ---- limiters: muscl_hll.hpp:33-57 ------------------------------------------------------
template <typename R>
__device__ __forceinline__ R minmod(R a, R b)
{
    const R zero(0);
    return (a * b <= zero) ? zero : ((dabs(a) < dabs(b)) ? a : b);
}

// fp32: minmod(a, b) is the median of a, b and 0 (same signs: the one of smaller magnitude;
// opposite signs or a zero: 0), one v_med3_f32. It differs from the select form only where
// a * b underflows to 0 with both of one sign (the select form gives 0) and in the NaN it
// propagates; the FP32 path is held to a relative L1 tolerance, the fp64 path keeps the form
// above.
__device__ __forceinline__ float minmod_med3(float a, float b)
{
    return __builtin_amdgcn_fmed3f(a, b, 0.0f);
}

template <int LIM, typename R>
__device__ __forceinline__ R limit_slope(R a, R b)
{
    if constexpr (sizeof(R) == 4) {
        if constexpr (LIM == kMinmod) {
            return minmod_med3(a, b);
        } else {
            return minmod_med3(0.5f * (a + b), minmod_med3(2.0f * a, 2.0f * b));
        }
    } else if constexpr (LIM == kMinmod) {
        return minmod(a, b);
    } else {
        const R half(0.5);
        const R two(2);
        return minmod(half * (a + b), minmod(two * a, two * b));
    }
}

// ---- mixed-precision hooks ----------------------------------------------------------------
template <typename R>
struct MixedOps;

template <>
struct MixedOps<double> {
    // cf^2 discriminant max(sum^2 - 4 cs2 van2, 0): muscl_hll.hpp:84
    static __device__ __forceinline__ double disc(double sum, double cs2, double van2)
    {
        return hmax(sum * sum - 4.0 * cs2 * van2, 0.0);
    }
    // p = (gamma-1) (E - ekin - emag) with ekin = 0.5 |m|^2 inv_rho: cpu_rows.hpp:146-153
    static __device__ __forceinline__ double pressure(double rho, double inv_rho, double mx,
                                                      double my, double mz, double e, double bx,
                                                      double by, double bz, double gm1)
    {
        (void)rho;
        const double ekin = 0.5 * (mx * mx + my * my + mz * mz) * inv_rho;
        const double emag = 0.5 * (bx * bx + by * by + bz * bz);
        return gm1 * (e - ekin - emag);
    }
};

template <>
struct MixedOps<float> {
    static __device__ __forceinline__ float disc(float sum, float cs2, float van2)
    {
        // fp32 x fp32 products are exact in fp64; one fp64 rounding for the difference.
        const double s = static_cast<double>(sum);
        const double d = s * s - 4.0 * (static_cast<double>(cs2) * static_cast<double>(van2));
        return static_cast<float>(d > 0.0 ? d : 0.0);
    }
    static __device__ __forceinline__ float pressure(float rho, float inv_rho, float mx, float my,
                                                     float mz, float e, float bx, float by,
                                                     float bz, float gm1)
    {
        const double dmx = mx, dmy = my, dmz = mz, dbx = bx, dby = by, dbz = bz;
        // 1/rho in fp64 from the fp32 reciprocal and one Newton step (relative error ~2^-46),
        // in place of an fp64 division
        const double r = rho;
        double inv = inv_rho;
        inv = __builtin_fma(__builtin_fma(-r, inv, 1.0), inv, inv);
        const double m2 = dmx * dmx + dmy * dmy + dmz * dmz;
        const double b2 = dbx * dbx + dby * dby + dbz * dbz;
        // e - m^2 / (2 rho) - B^2 / 2, each subtraction fused with its product
        const double eint =
            __builtin_fma(-0.5, b2, __builtin_fma(-0.5 * m2, inv, static_cast<double>(e)));
        return static_cast<float>(static_cast<double>(gm1) * eint);
    }
};

// ---- fast magnetosonic speed along DIR: muscl_hll.hpp:70-87 --------------------------------
template <int DIR, typename R>
__device__ __forceinline__ R fast_speed(const R (&W)[kNC], R gamma)
{
    using D = FaceDir<DIR>;
    const R half(0.5);
    const R inv_rho = drcp(W[RHO]);
    const R cs2 = gamma * W[PR] * inv_rho;
    const R bn2 = W[D::BN] * W[D::BN];
    const R b2 = bn2 + W[D::BT] * W[D::BT] + W[BZ] * W[BZ];
    const R van2 = bn2 * inv_rho;
    const R sum = fmadd(b2, inv_rho, cs2);  // cs2 + va2, va2 = B^2 / rho
    const R disc = MixedOps<R>::disc(sum, cs2, van2);
    const R cf2 = half * (sum + dsqrt(disc));
    return dsqrt(hmax(cf2, R(0)));
}

// ---- conserved state and GLM-MHD flux along DIR: muscl_hll.hpp:89-124 ---------------------
template <int DIR, typename R>
__device__ __forceinline__ void cons_flux(const R (&W)[kNC], R inv_gm1, R ch2, R (&U)[kNC],
                                          R (&F)[kNC])
{
    using D = FaceDir<DIR>;
    const R half(0.5);
    const R rho = W[RHO], vn = W[D::N], vt = W[D::T], vz = W[VZ], p = W[PR];
    const R bn = W[D::BN], bt = W[D::BT], bz = W[BZ], psi = W[PSI];
    const R mn = rho * vn;
    const R mt = rho * vt;
    const R mz = rho * vz;
    const R v2 = vn * vn + vt * vt + vz * vz;
    const R b2 = bn * bn + bt * bt + bz * bz;
    // pmag = B^2 / 2 enters both sums below
    const R energy = fmadd(half, b2, p * inv_gm1 + half * rho * v2);
    const R pstar = fmadd(half, b2, p);
    const R vb = vn * bn + vt * bt + vz * bz;

    U[RHO] = rho;
    U[D::N] = mn;
    U[D::T] = mt;
    U[VZ] = mz;
    U[PR] = energy;
    U[D::BN] = bn;
    U[D::BT] = bt;
    U[BZ] = bz;
    U[PSI] = psi;

    F[RHO] = mn;
    F[D::N] = mn * vn + pstar - bn * bn;
    F[D::T] = mn * vt - bn * bt;
    F[VZ] = mn * vz - bn * bz;
    F[PR] = (energy + pstar) * vn - bn * vb + psi * bn;
    F[D::BN] = psi;
    F[D::BT] = vn * bt - vt * bn;
    F[BZ] = vn * bz - vz * bn;
    F[PSI] = ch2 * bn;
}

// ---- HLL + GLM flux from reconstructed face states: muscl_hll.hpp:145-209 ----------------
// WL/WR are the MUSCL states; Wm/W0 are the cells left/right of the face, used when either
// reconstructed state is non-positive or non-finite (both sides then revert to first order).
// 0 < x < inf (false for NaN): one class test for positive denormal or positive normal
__device__ __forceinline__ bool positive_finite(double x)
{
    return __builtin_amdgcn_class(x, (1 << 7) | (1 << 8));
}
__device__ __forceinline__ bool positive_finite(float x)
{
    return __builtin_amdgcn_classf(x, (1 << 7) | (1 << 8));
}

// True on every lane when `pred` holds on some lane of `live`.
__device__ __forceinline__ bool any_lane(bool pred, unsigned long long live)
{
    return (__builtin_amdgcn_ballot_w64(pred) & live) != 0;
}

// `live`: the lanes whose flux is used; the rare-path branches below are taken for the whole
// wave when any live lane needs them.
template <int DIR, typename R>
__device__ __forceinline__ void hll_face(R (&WL)[kNC], R (&WR)[kNC], const R (&Wm)[kNC],
                                         const R (&W0)[kNC], R gamma, R inv_gm1, R c_h,
                                         R (&F)[kNC], unsigned long long live)
{
    using D = FaceDir<DIR>;
    const R zero(0);
    const bool ok = positive_finite(WL[RHO]) && positive_finite(WL[PR]) &&
                    positive_finite(WR[RHO]) && positive_finite(WR[PR]);
    // the fallback is rare: skip its selects for the whole wave unless some lane needs it
    if (any_lane(!ok, live)) {
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            WL[c] = ok ? WL[c] : Wm[c];
            WR[c] = ok ? WR[c] : W0[c];
        }
    }
    const R cfL = fast_speed<DIR>(WL, gamma);
    const R cfR = fast_speed<DIR>(WR, gamma);
    const R sL = hmin(hmin(WL[D::N] - cfL, WR[D::N] - cfR), -c_h);
    const R sR = hmax(hmax(WL[D::N] + cfL, WR[D::N] + cfR), c_h);
    const R ch2 = c_h * c_h;

    R UL[kNC], FL[kNC], UR[kNC], FR[kNC];
    cons_flux<DIR>(WL, inv_gm1, ch2, UL, FL);
    cons_flux<DIR>(WR, inv_gm1, ch2, UR, FR);

    const R half(0.5);
    const bool use_left = sL >= zero;
    const bool use_right = sR <= zero;
    if constexpr (sizeof(R) == 4) {
        // fp32: with sL clamped to <= 0 and sR to >= the smallest normal float, the HLL formula
        // itself returns FL (sL >= 0, also when both speeds vanish) or FR (sR <= 0), to
        // rounding, so no per-component selects are needed. A span below 1e-14 with
        // sL < 0 < sR (every signal speed of the grid below 5e-15) takes the reference's
        // average (FL + FR) / 2 through the same formula, with weights 1/2, 1/2 and no
        // dissipation term.
        const bool degenerate = !use_left && !use_right && (sR - sL) < Lit<R>::kDegenerate;
        // the degenerate lanes weigh FL and FR by 1 / 2 each (speeds -1 and 1) and drop the
        // dissipation term
        const R sLc = degenerate ? R(-1) : hmin(sL, zero);
        const R sRc = degenerate ? R(1) : hmax(sR, R(1.17549435e-38f));
        const R inv_denom = drcp(sRc - sLc);
        const R a = sRc * inv_denom;
        const R b = -sLc * inv_denom;
        const R k = degenerate ? zero : sLc * sRc * inv_denom;
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            F[c] = a * FL[c] + (b * FR[c] + k * (UR[c] - UL[c]));
        }
        return;
    }
    const R denom = sR - sL;
    const R inv_denom = drcp(denom);
    const R sLsR = sL * sR;
    const bool degenerate = dabs(denom) < Lit<R>::kDegenerate;
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        const R hll = (sR * FL[c] - sL * FR[c] + sLsR * (UR[c] - UL[c])) * inv_denom;
        F[c] = use_left ? FL[c] : (use_right ? FR[c] : hll);
    }
    // sR - sL spans at least 2 c_h, so the degenerate average is rare: wave-uniform skip
    if (any_lane(degenerate && !use_left && !use_right, live)) {
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            const R mid = half * (FL[c] + FR[c]);
            F[c] = (degenerate && !use_left && !use_right) ? mid : F[c];
        }
    }
}

// ---- conserved -> primitive: cpu_rows.hpp:130-155 ------------------------------------------
template <typename R>
__device__ __forceinline__ void to_prim(const R (&U)[kNC], R gm1, R (&W)[kNC])
{
    const R inv_rho = drcp(U[RHO]);
    W[RHO] = U[RHO];
    W[VX] = U[1] * inv_rho;
    W[VY] = U[2] * inv_rho;
    W[VZ] = U[3] * inv_rho;
    W[PR] = MixedOps<R>::pressure(U[RHO], inv_rho, U[1], U[2], U[3], U[4], U[5], U[6], U[7], gm1);
    W[BX] = U[5];
    W[BY] = U[6];
    W[BZ] = U[7];
    W[PSI] = U[8];
}

// ---- CFL signal speed max(|vx|+cf_x, |vy|+cf_y) of one cell: muscl_hll.hpp:211-239 ---------
template <typename R>
__device__ __forceinline__ R cell_max_speed(const R (&U)[kNC], R gamma, R gm1)
{
    const R inv_rho = drcp(U[RHO]);
    R W[kNC];
    W[RHO] = U[RHO];
    W[VX] = U[1] * inv_rho;
    W[VY] = U[2] * inv_rho;
    W[VZ] = U[3] * inv_rho;
    W[PR] = MixedOps<R>::pressure(U[RHO], inv_rho, U[1], U[2], U[3], U[4], U[5], U[6], U[7], gm1);
    W[BX] = U[5];
    W[BY] = U[6];
    W[BZ] = U[7];
    W[PSI] = R(0);
    // x: (n, t) = (x, y); y: (n, t) = (y, x) -- the y sum order is by^2 + bx^2 + bz^2
    const R sx = dabs(W[VX]) + fast_speed<0>(W, gamma);
    const R sy = dabs(W[VY]) + fast_speed<1>(W, gamma);
    return hmax(sx, sy);
}

// ---- non-ideal electric field: nonideal.hpp:63-102 -----------------------------------------
struct NonIdeal {
    int ohmic, hall, ambipolar;
    double eta_ohm, eta_hall, eta_ambipolar;
};

template <typename R>
__device__ __forceinline__ void nonideal_e(const R (&J)[3], const R (&B)[3], const NonIdeal& cfg,
                                           R (&E)[3])
{
    E[0] = R(0);
    E[1] = R(0);
    E[2] = R(0);
    if (cfg.ohmic) {
        const R eta = static_cast<R>(cfg.eta_ohm);
        E[0] += eta * J[0];
        E[1] += eta * J[1];
        E[2] += eta * J[2];
    }
    if (!(cfg.hall || cfg.ambipolar)) {
        return;
    }
    const R B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    if (B2 < R(1.0e-14)) {
        return;
    }
    const R jb0 = J[1] * B[2] - J[2] * B[1];
    const R jb1 = J[2] * B[0] - J[0] * B[2];
    const R jb2 = J[0] * B[1] - J[1] * B[0];
    if (cfg.hall) {
        const R eta = static_cast<R>(cfg.eta_hall);
        const R invB = drcp(dsqrt(B2));
        E[0] += eta * jb0 * invB;
        E[1] += eta * jb1 * invB;
        E[2] += eta * jb2 * invB;
    }
    if (cfg.ambipolar) {
        const R eta = static_cast<R>(cfg.eta_ambipolar);
        const R a0 = jb1 * B[2] - jb2 * B[1];
        const R a1 = jb2 * B[0] - jb0 * B[2];
        const R a2 = jb0 * B[1] - jb1 * B[0];
        const R invB2 = drcp(B2);
        E[0] += eta * a0 * invB2;
        E[1] += eta * a1 * invB2;
        E[2] += eta * a2 * invB2;
    }
}

// ---- ghost index maps: boundary_conditions.hpp:26-30 (periodic), :44-46 (outflow) ---------
// Maps a global index g of an axis with n interior cells and ng ghosts to the interior cell
// whose value the ghost holds after the x-then-y ghost fill. Interior indices map to themselves.
__device__ __host__ __forceinline__ int ghost_source(int g, int n, int ng, int bc)
{
    if (g >= ng && g < ng + n) {
        return g;
    }
    if (bc == 0) {  // periodic: ng + ((g - ng) mod n)
        const int off = g - ng;
        return ng + ((off % n) + n) % n;
    }
    return g < ng ? ng : ng + n - 1;  // outflow: nearest interior cell
}

}  // namespace mhdhip
