// Per-cell arithmetic of the filtration step, written once for the HIP
// kernels and for a host build of the same code (used to check the kernels'
// arithmetic against the oracle on the CPU).
//
// Parity rule: every function below produces the same IEEE bits as the
// reference CPU path (cpu.cpp, Two-phase/two-phase.cpp,
// Three-phase/three-phase.cpp, gauss.cpp, energy.cpp) for finite operands in
// the normal range, up to the sign of an exact zero.  Four transformations
// are used, each exact:
//   * a / b with a constant b is exact_div(a, b, RN(1/b)): one FMA residual
//     correction of a*RN(1/b) returns RN(a/b) (Markstein's theorem, Muller et
//     al., Handbook of Floating-Point Arithmetic, 2nd ed., Thm 4.10);
//   * x / 2. is x * 0.5 (a power-of-two scaling is exact);
//   * the upwind parts (x + |x|) / 2 and (x - |x|) / 2 are max(x, 0) and
//     min(x, 0) (for |x| < 2^1023: x +- |x| is 2x or +-0, exactly);
//   * terms that are exact zeros in the reference (the CPU path's capillary
//     pressures and well rates, three-phase.cpp:23-143,529-550) are folded:
//     x + 0 == x and x - (+-0) == x for x != 0.
// The code is built with -ffp-contract=off so no other FMA is formed.
#pragma once

#include <math.h>
#include <stdint.h>

#include "capaz_params.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CAPAZ_HD __host__ __device__ __forceinline__
#else
#define CAPAZ_HD inline
#endif

namespace capaz {

// ---- arithmetic --------------------------------------------------------------

// RN(a / b) given rb = RN(1 / b): q = RN(a*rb) is within one ulp of a/b, the
// residual a - b*q is exact in an FMA, and one correction rounds correctly.
CAPAZ_HD double exact_div(double a, double b, double rb) {
    double q = a * rb;
    double r = fma(-q, b, a);
    return fma(r, rb, q);
}

// RN(1/b): the hardware reciprocal estimate, two Newton-Raphson steps and a
// final residual correction (Markstein: the correction of an estimate within
// one ulp of 1/b rounds to RN(1/b)).  Operands outside [2^-1000, 2^1000] in
// magnitude, zeros and non-finite values take the IEEE division, out of line
// so that the common path is not flattened with it.  Checked against 1.0/b
// on the device by tests/test_selftest.py.
#if defined(__HIPCC__)
__device__ __attribute__((noinline)) inline double rcp_ieee(double b) { return 1. / b; }
#endif

CAPAZ_HD double rcp_rn(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double ab = fabs(b);
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.);
    y = fma(e, y, y);
    e = fma(-b, y, 1.);
    y = fma(e, y, y);
    e = fma(-b, y, 1.);
    y = fma(e, y, y);
    if (__builtin_expect(!(ab >= 0x1p-1000 && ab <= 0x1p1000), 0)) y = rcp_ieee(b);
    return y;
#else
    return 1. / b;
#endif
}

// rcp_rn for a pivot the caller has checked to satisfy |b| >= 1e-8: only the
// upper end of the range (and NaN) can take the IEEE division.
CAPAZ_HD double rcp_rn_pivot(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double ab = fabs(b);
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.);
    y = fma(e, y, y);
    e = fma(-b, y, 1.);
    y = fma(e, y, y);
    e = fma(-b, y, 1.);
    y = fma(e, y, y);
    if (__builtin_expect(!(ab <= 0x1p1000), 0)) y = rcp_ieee(b);
    return y;
#else
    return 1. / b;
#endif
}

CAPAZ_HD double half(double x) { return x * 0.5; }
// x clipped to [0, 1]; a NaN passes through (so a non-finite state spreads
// through k_r as in the oracle; fmin/fmax would turn it into 0).
CAPAZ_HD double clamp01(double x) { return x < 0. ? 0. : (x > 1. ? 1. : x); }

// ---- grid predicates ------------------------------------------------------------

// assign_roS_nr's interior test (cpu.cpp:621), shared by the three-phase
// assign_P_Xi/Newton (three-phase.cpp:225,288) and energy.cpp:237,272,330,372.
CAPAZ_HD bool interior3(const DevParams& p, int i, int j, int k) {
    return (((i != 0) && (i != p.Nx - 1)) || (p.Nx < 2)) && (j != 0) && (j != p.Ny - 1) &&
           (((k != 0) && (k != p.Nz - 1)) || (p.Nz < 2));
}

// The two-phase Newton's interior test (two-phase.cpp:131).
CAPAZ_HD bool interior_newton2(const DevParams& p, int i, int j, int k) {
    return (i != 0) && (i != p.Nx - 1) && (j != 0) && (j != p.Ny - 1) &&
           (((k != 0) && (k != p.Nz - 1)) || (p.Nz < 2));
}

// Border_S/Border_P's boundary test (two-phase.cpp:165; three-phase.cpp:347).
CAPAZ_HD bool is_boundary(const DevParams& p, int i, int j, int k) {
    return p.model == 2 ? !interior_newton2(p, i, j, k) : !interior3(p, i, j, k);
}

// set_boundary_basic_coordinate (cpu.cpp:107-141).
CAPAZ_HD int64_t inward(const DevParams& p, int i, int j, int k) {
    int i1 = i, j1 = j, k1 = k;
    if (i == 0 && p.Nx > 2) i1++;
    if (i == p.Nx - 1 && p.Nx > 2) i1--;
    if (j == 0) j1++;
    if (j == p.Ny - 1) j1--;
    if (k == 0 && p.Nz > 2) k1++;
    if (k == p.Nz - 1 && p.Nz > 2) k1--;
    return (int64_t)i1 + (int64_t)j1 * p.sy + (int64_t)k1 * p.sz;
}

// ---- closures ---------------------------------------------------------------------

// x^y in the reference closures (the van Genuchten and Brooks-Corey k_r, the
// energy model's mu_g(T) and lambda_g(T), a Corey exponent read at run time).
// On the device: exp(y * log|x|) from the device library's log and exp, about
// 130 VALU instructions against some 215 for its pow, with the log computed
// once where powers share a base (pow_base / pow_of).  The roundings of log x
// and of y * log x put up to 2 |y ln x| ulp into the result (tens of ulp for
// the closures' bases, x >= 1e-15), within the tolerance the reference
// closures are held to, since the device library's pow is not glibc's either.  pow's other
// cases, which the closures reach only for states outside their range, by
// selects (no call, which would cost registers in every kernel that holds a
// closure's code): a zero exponent gives 1, a negative finite base gives
// +-|x|^y for an integer exponent and NaN otherwise.  The host build keeps
// glibc's pow, whose bits the reference's golden vectors hold.
#if defined(__HIP_DEVICE_COMPILE__)
// ln x for x > 0 (pow_base's log|x|): x = 2^k m with m in [sqrt(1/2),
// sqrt(2)) by the hardware's frexp (subnormals included), then
// ln m = 2 atanh(s), s = (m - 1) / (m + 1) (|s| <= 0.1716; m - 1 is exact),
// by the series 2 s (1 + s^2/3 + s^4/5 + ... + s^20/21), whose next term is
// below 1e-17 of the sum, and k ln 2 with ln 2 split so that k * LN2_HI is
// exact.  s carries ~2^-52 relative error (the rounded m + 1 and a reciprocal
// with two Newton steps), which the result keeps for x near 1; for other x
// the final additions' rounding dominates, as in a correctly rounded log.
// 0 gives -inf, +inf gives +inf, NaN and negative x give NaN.
__device__ __forceinline__ double log_cl(double x) {
    constexpr double kLn2Hi = 0x1.62e42fefa3800p-1, kLn2Lo = 0x1.ef35793c76730p-45;
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1) for finite x > 0
    int k = __builtin_amdgcn_frexp_exp(x);
    const bool low = m < 0x1.6a09e667f3bcdp-1;  // sqrt(1/2)
    m = low ? 2. * m : m;
    k = low ? k - 1 : k;
    const double q = m + 1.;
    double r = __builtin_amdgcn_rcp(q);
    r = fma(fma(-q, r, 1.), r, r);
    r = fma(fma(-q, r, 1.), r, r);
    const double s = (m - 1.) * r, z = s * s;
    double poly = 1. / 21.;
    poly = fma(poly, z, 1. / 19.);
    poly = fma(poly, z, 1. / 17.);
    poly = fma(poly, z, 1. / 15.);
    poly = fma(poly, z, 1. / 13.);
    poly = fma(poly, z, 1. / 11.);
    poly = fma(poly, z, 1. / 9.);
    poly = fma(poly, z, 1. / 7.);
    poly = fma(poly, z, 1. / 5.);
    poly = fma(poly, z, 1. / 3.);
    const double lm = 2. * fma(s * z, poly, s);  // ln m
    const double kd = (double)k;
    double l = fma(kd, kLn2Hi, fma(kd, kLn2Lo, lm));
    l = x == 0. ? -__builtin_inf() : l;
    l = x == __builtin_inf() ? __builtin_inf() : l;
    return x > 0. || x == 0. ? l : __builtin_nan("");  // NaN and x < 0: NaN
}
#endif

struct PowBase {
    double x, lx;  // the base and (device) log|x|
};
CAPAZ_HD PowBase pow_base(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return {x, log_cl(fabs(x))};
#else
    return {x, 0.};
#endif
}
CAPAZ_HD double pow_of(const PowBase& b, double y) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = exp(y * b.lx);
    const bool integer = y == trunc(y), odd = integer && 0.5 * y != trunc(0.5 * y);
    r = b.x < 0. ? (integer ? (odd ? -r : r) : __builtin_nan("")) : r;
    return y == 0. ? 1. : r;
#else
    return pow(b.x, y);
#endif
}
CAPAZ_HD double pow_cl(double x, double y) { return pow_of(pow_base(x), y); }

CAPAZ_HD double corey_kr(const DevParams& p, double s) {
    s = clamp01(s);
    if (p.corey_n == 2.) return s * s;
    return pow_cl(s, p.corey_n);
}

// pow(x, 0.5) and pow(t, 2.) of the van Genuchten closures.  On the device:
// the correctly rounded sqrt(x) (+ 0. turns sqrt(-0) = -0 into pow's +0) and
// t * t, some 180 VALU instructions fewer each than the device library's pow,
// whose bits are not glibc's either.  The host build keeps glibc's pow, whose
// bits the reference's golden vectors hold (glibc's pow(x, 0.5) differs from
// sqrt(x) in ~1 of 1000 draws, by one ulp).  Device results checked against
// the correctly rounded operations by tests/test_selftest.py.
CAPAZ_HD double pow_half(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return sqrt(x) + 0.;
#else
    return pow(x, 0.5);
#endif
}
CAPAZ_HD double pow_two(double t) {
#if defined(__HIP_DEVICE_COMPILE__)
    return t * t;
#else
    return pow(t, 2.);
#endif
}

// van Genuchten-type k_r (three-phase.cpp:146-186); p.vg_a = A/(A-1.),
// p.vg_b = (A-1.)/A, p.vg_c = 2.*(A-1.)/A.
// pSwe = pow(Swe, p.vg_a), shared by k_w and k_n (read only where they use
// it), and b1 = pow_base(1. - pSwe), the base of their pow(1. - pSwe, .).
CAPAZ_HD double vg_k_w(const DevParams& p, double Swe, const PowBase& b1) {
    if (Swe >= 1e-3) {
        double t = 1. - pow_of(b1, p.vg_b);
        return pow_half(Swe) * pow_two(t);
    }
    return 0;
}
CAPAZ_HD double vg_k_g(const DevParams& p, double Sge) {
    if (Sge >= 1e-3) return pow_half(Sge) * pow_cl(1. - pow_cl(1. - Sge, p.vg_a), p.vg_c);
    return 0;
}
CAPAZ_HD double vg_k_n(const DevParams& p, double Swe, double Sne, const PowBase& b1) {
    double Sge = 1. - Swe - Sne;
    if (Sne >= 1e-3) {
        double knw = pow_half(1. - Swe) * pow_of(b1, p.vg_c);
        double kng = pow_half(Sne) * pow_two(1. - pow_cl(1. - pow_cl(Sne, p.vg_a), p.vg_b));
        return Sne * knw * kng / (1 - Swe) / (1 - Sge);
    }
    return 0;
}

// Three-phase relative permeabilities k_w, k_n, k_g of a cell
// (three-phase.cpp:202-208).
CAPAZ_HD void relperm3(const DevParams& p, double S_w, double S_n, double& kw, double& kn, double& kg) {
    double Swe = exact_div(S_w - p.S_wr, p.den3, p.r_den3);
    double Sne = exact_div(S_n - p.S_nr, p.den3, p.r_den3);
    double Sge = 1. - Swe - Sne;
    if (p.closure == 1) {
        kw = corey_kr(p, Swe);
        kn = corey_kr(p, Sne);
        kg = corey_kr(p, Sge);
    } else {
        const double pSwe = Swe >= 1e-3 || Sne >= 1e-3 ? pow_cl(Swe, p.vg_a) : 0.;
        const PowBase b1 = pow_base(1. - pSwe);
        kw = vg_k_w(p, Swe, b1);
        kg = vg_k_g(p, Sge);
        kn = vg_k_n(p, Swe, Sne, b1);
    }
}

// Two-phase k_w, k_n and P_k (two-phase.cpp:6-12; Corey: no P_k).
CAPAZ_HD void relperm2(const DevParams& p, double S_n, double& kw, double& kn, double& Pk) {
    if (p.closure == 1) {
        double se = exact_div(1. - S_n - p.S_wr, p.den2c, p.r_den2c);
        kw = corey_kr(p, se);
        kn = corey_kr(p, 1. - se);
        Pk = 0.;
    } else {
        double se = exact_div(1. - S_n - p.S_wr, p.den2, p.r_den2);
        const PowBase b = pow_base(se);
        kw = pow_of(b, p.bc_ew);
        kn = (1. - se) * (1. - se) * (1 - pow_of(b, p.bc_en));
        Pk = p.P_d * pow_of(b, p.bc_epk);
    }
}

// Energy-model viscosities mu(T) (three-phase.cpp:212-214).
CAPAZ_HD void mu_of_T(double T, double& mw, double& mn, double& mg) {
    mw = 1. / (29.21 * T - 7506.64);
    mn = 7.256E-10 * exp(4141.9 / T);
    mg = 1.717E-5 * pow_cl(exact_div(T, 273., 1. / 273.), 0.683);
}

// Xi = -1. * K * k / mu (two-phase.cpp:16-17, three-phase.cpp:216-222).
CAPAZ_HD double xi_const(double K, double kr, double mu, double r_mu) { return exact_div((-1.) * K * kr, mu, r_mu); }
CAPAZ_HD double xi_var(double K, double kr, double mu) { return (-1.) * K * kr / mu; }

// Densities (cpu.cpp:143-169).
CAPAZ_HD double ro_w_of(const DevParams& p, double P) { return p.ro0_w * (1. + p.beta_w * (P - p.P_atm)); }
CAPAZ_HD double ro_n_of(const DevParams& p, double P) { return p.ro0_n * (1. + p.beta_n * (P - p.P_atm)); }
CAPAZ_HD double ro_g_of(const DevParams& p, double P) { return exact_div(p.ro0_g * P, p.P_atm, p.r_P_atm); }

// Energy-model densities (cpu.cpp:150-156, energy.cpp:146-170); T_0 = 273.
constexpr double kT0 = 273.;
constexpr double kAlfaW = 1.32E-7, kAlfaN = 9.2E-7;
CAPAZ_HD double ro_w_T(const DevParams& p, double P, double T) {
    return p.ro0_w * (1. + p.beta_w * (P - p.P_atm) - kAlfaW * (T - kT0));
}
CAPAZ_HD double ro_n_T(const DevParams& p, double P, double T) {
    return p.ro0_n * (1. + p.beta_n * (P - p.P_atm) - kAlfaN * (T - kT0));
}
CAPAZ_HD double ro_g_T(const DevParams& p, double P, double T) {
    return p.ro0_g * exact_div(P, p.P_atm, p.r_P_atm) * (kT0 / T);
}

// ---- Darcy flux (cpu.cpp:236-264, 619-731) -----------------------------------------

// directed_difference: x1, x2 are the left/right driving forces, (Xi, ro) of
// the cell (c) and of its minus (m) and plus (p) neighbours.  The
// reference's upwind parts (x + |x|) / 2 and (x - |x|) / 2 are max(x, 0) and
// min(x, 0): for |x| < 2^1023, x +- |x| is 2x (exact) or +-0 and the halving
// is exact, so they agree up to the sign of a zero.  A NaN x (a NaN pressure
// at c or at the neighbour) gives NaN parts in the reference and 0 here; the
// flux is NaN either way, through the density of the same cell (ro_c, or
// ro_m / ro_p times the 0), a function of that pressure.
CAPAZ_HD double pos_part(double x) { return fmax(x, 0.); }
CAPAZ_HD double neg_part(double x) { return fmin(x, 0.); }
CAPAZ_HD double directed(double x1, double x2, double Xic, double roc, double Xim, double rom, double Xip, double rop,
                         double h, double rh) {
    double t1 = (pos_part(x2) - neg_part(x1)) * (-1.) * Xic * roc;
    double t2 = pos_part(x1) * (-1.) * Xim * rom;
    double t3 = neg_part(x2) * (-1.) * Xip * rop;
    return exact_div(t1 - t2 + t3, h, rh);
}

// directed() from its minus face's parts computed beforehand: nx1 =
// neg_part(x1) and t2 = directed_t2(x1, Xim, rom); the same operations in
// the same order, so the same result.
CAPAZ_HD double directed_t2(double x1, double Xim, double rom) { return pos_part(x1) * (-1.) * Xim * rom; }
CAPAZ_HD double directed_parts(double nx1, double t2, double x2, double Xic, double roc, double Xip, double rop,
                               double h, double rh) {
    double t1 = (pos_part(x2) - nx1) * (-1.) * Xic * roc;
    double t3 = neg_part(x2) * (-1.) * Xip * rop;
    return exact_div(t1 - t2 + t3, h, rh);
}

// The upwind flux of one phase along one axis, from the face differences
// dm = (P_c - P_m)/h and dp = (P_p - P_c)/h; in y the driving force carries
// + g*ro_c (cpu.cpp:699-705).
CAPAZ_HD double flux_x(double dm, double dp, double Xic, double roc, double Xim, double rom, double Xip, double rop,
                       double h, double rh) {
    return directed(-1. * dm, -1. * dp, Xic, roc, Xim, rom, Xip, rop, h, rh);
}
CAPAZ_HD double flux_y(const DevParams& p, double dm, double dp, double Xic, double roc, double Xim, double rom,
                       double Xip, double rop) {
    double gr = p.g * roc;
    return directed(-1. * dm + gr, -1. * dp + gr, Xic, roc, Xim, rom, Xip, rop, p.hy, p.r_hy);
}

// ---- dense solve (gauss.cpp:6-104) -------------------------------------------------------

// reverse_matrix on a row-major n x n matrix held in registers.  Same pivot
// rule as the reference: pivot only if |a_kk| < 1e-8, then swap with every
// later row whose (signed) entry exceeds 1e-8; if none, return -1 and leave
// `a` as it was.  Row k is divided by the pivot through one correctly rounded
// reciprocal and exact_div per entry.
template <int N>
CAPAZ_HD int reverse_matrix(double* a) {
    constexpr int N1 = 2 * N;
    const double e = 1E-8;
    double m[N][N1];
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int j = 0; j < N; j++) m[i][j] = a[j + N * i];
#pragma unroll
        for (int j = N; j < N1; j++) m[i][j] = (j == N + i) ? 1. : 0.;
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
        double f = m[k][k];
        if (fabs(f) < e) {
            bool found = false;
#pragma unroll
            for (int i = k + 1; i < N; i++) {
                if (m[i][k] > e) {
                    found = true;
#pragma unroll
                    for (int j = 0; j < N1; j++) {
                        double t = m[k][j];
                        m[k][j] = m[i][j];
                        m[i][j] = t;
                    }
                    f = m[k][k];
                }
            }
            if (!found) return -1;
        }
        double rf = 1. / f;
#pragma unroll
        for (int j = 0; j < N1; j++) m[k][j] = exact_div(m[k][j], f, rf);
#pragma unroll
        for (int i = 0; i < N; i++) {
            if (i != k) {
                double f1 = m[i][k];
#pragma unroll
                for (int j = 0; j < N1; j++) m[i][j] -= (m[k][j] * f1);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) a[j + N * i] = m[i][j + N];
    return 0;
}

// reverse_matrix<3> when no pivot search is needed: the same operations on
// every entry that is not known to be exactly 0 or 1, so the result has the
// same bits (entries known to be zero only lose the sign of a zero, which
// never reaches a non-zero result).  Returns false, leaving `a` untouched,
// if some pivot has |f| < 1e-8; the caller then runs the general routine.
CAPAZ_HD bool reverse3_nopivot(double* a) {
    const double e = 1E-8;
    // k = 0: row 0 = [1, a01/f, a02/f | 1/f, 0, 0]
    double f = a[0];
    if (fabs(f) < e) return false;
    double rf = rcp_rn_pivot(f);
    double r01 = exact_div(a[1], f, rf), r02 = exact_div(a[2], f, rf), r03 = rf;
    // rows 1, 2 lose column 0: [0, b_i1, b_i2 | b_i3, delta_i1, delta_i2]
    double f1 = a[3];
    double b11 = a[4] - r01 * f1, b12 = a[5] - r02 * f1, b13 = 0. - r03 * f1;
    double f2 = a[6];
    double b21 = a[7] - r01 * f2, b22 = a[8] - r02 * f2, b23 = 0. - r03 * f2;
    // k = 1: row 1 = [0, 1, b12/f | b13/f, 1/f, 0]
    f = b11;
    if (fabs(f) < e) return false;
    rf = rcp_rn_pivot(f);
    double s12 = exact_div(b12, f, rf), s13 = exact_div(b13, f, rf), s14 = rf;
    f1 = r01;  // row 0
    double c02 = r02 - s12 * f1, c03 = r03 - s13 * f1, c04 = 0. - s14 * f1;
    f2 = b21;  // row 2
    double c22 = b22 - s12 * f2, c23 = b23 - s13 * f2, c24 = 0. - s14 * f2;
    // k = 2: row 2 = [0, 0, 1 | c23/f, c24/f, 1/f]
    f = c22;
    if (fabs(f) < e) return false;
    rf = rcp_rn_pivot(f);
    double t23 = exact_div(c23, f, rf), t24 = exact_div(c24, f, rf), t25 = rf;
    f1 = c02;  // row 0
    double o03 = c03 - t23 * f1, o04 = c04 - t24 * f1, o05 = 0. - t25 * f1;
    f2 = s12;  // row 1
    double o13 = s13 - t23 * f2, o14 = s14 - t24 * f2, o15 = 0. - t25 * f2;
    a[0] = o03; a[1] = o04; a[2] = o05;
    a[3] = o13; a[4] = o14; a[5] = o15;
    a[6] = t23; a[7] = t24; a[8] = t25;
    return true;
}

// The general routine, kept out of line on the device: it runs only when a
// pivot is below 1e-8, and inlining it would hold its 3 x 6 matrix in
// registers across the whole Newton loop.
#if defined(__HIPCC__)
__host__ __device__ __attribute__((noinline)) inline int reverse_matrix3_general(double* a) {
    return reverse_matrix<3>(a);
}
#else
inline int reverse_matrix3_general(double* a) { return reverse_matrix<3>(a); }
#endif

// reverse3_nopivot for a matrix whose a[2] is +0 (the three-phase Jacobian's
// dF[2], three-phase.cpp:321): a02/f is +-0, so b12 = a12 - (+-0)*f1 = a12,
// b22 = a22 and c02 = (+-0) - s12*r01 = -(s12*r01), all exact.  The
// eliminations against a zero of the identity block, 0. - x, are written
// -(x): the same value up to the sign of an exact zero, which no non-zero
// entry of the inverse or of the Newton update depends on.
CAPAZ_HD bool reverse3_nopivot_a02zero(double* a) {
    const double e = 1E-8;
    double f = a[0];
    if (fabs(f) < e) return false;
    double rf = rcp_rn_pivot(f);
    double r01 = exact_div(a[1], f, rf), r03 = rf;
    double f1 = a[3];
    double b11 = a[4] - r01 * f1, b12 = a[5], b13 = -(r03 * f1);
    double f2 = a[6];
    double b21 = a[7] - r01 * f2, b22 = a[8], b23 = -(r03 * f2);
    f = b11;
    if (fabs(f) < e) return false;
    rf = rcp_rn_pivot(f);
    double s12 = exact_div(b12, f, rf), s13 = exact_div(b13, f, rf), s14 = rf;
    f1 = r01;
    double c02 = -(s12 * f1), c03 = r03 - s13 * f1, c04 = -(s14 * f1);
    f2 = b21;
    double c22 = b22 - s12 * f2, c23 = b23 - s13 * f2, c24 = -(s14 * f2);
    f = c22;
    if (fabs(f) < e) return false;
    rf = rcp_rn_pivot(f);
    double t23 = exact_div(c23, f, rf), t24 = exact_div(c24, f, rf), t25 = rf;
    f1 = c02;
    double o03 = c03 - t23 * f1, o04 = c04 - t24 * f1, o05 = -(t25 * f1);
    f2 = s12;
    double o13 = s13 - t23 * f2, o14 = s14 - t24 * f2, o15 = -(t25 * f2);
    a[0] = o03; a[1] = o04; a[2] = o05;
    a[3] = o13; a[4] = o14; a[5] = o15;
    a[6] = t23; a[7] = t24; a[8] = t25;
    return true;
}

// reverse_matrix<3> with the no-pivot path first.  `a` stays in registers;
// only the rare fallback copies it to a stack buffer for the out-of-line call
// (reverse3_nopivot leaves `a` unchanged when it fails).
template <bool A02ZERO = false>
CAPAZ_HD int reverse3(double* a) {
    if (A02ZERO ? reverse3_nopivot_a02zero(a) : reverse3_nopivot(a)) return 0;
    double buf[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) buf[i] = a[i];
    const int rc = reverse_matrix3_general(buf);
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i] = buf[i];
    return rc;
}

// ---- Newton solves ------------------------------------------------------------------------

// Two-phase fixed-iteration 2x2 Newton by Cramer's rule (two-phase.cpp:25-56).
// CL: the fused kernels' closure id (0 the reference's Brooks-Corey, 1 and 2
// Corey), or -1 for p.closure at run time.
template <int CL = -1>
CAPAZ_HD void newton2(const DevParams& p, double& P, double& Sn, double roS_w, double roS_n) {
    for (int w = 1; w <= p.iterations; w++) {
        double P_k, PkS;
        if (CL >= 1 || (CL < 0 && p.closure == 1)) {
            P_k = 0.;
            PkS = 0.;
        } else {
            double S_e = exact_div(1 - Sn - p.S_wr, p.den2, p.r_den2);
            const PowBase b = pow_base(S_e);
            P_k = p.P_d * pow_of(b, p.bc_epk);
            double AAA = pow_of(b, p.bc_eaaa);
            PkS = exact_div(AAA * p.P_d, p.bc_pks_den, p.r_bc_pks_den);
        }
        double F1 = p.ro0_w * (1. + p.beta_w * (P - p.P_atm)) * (1. - Sn) - roS_w;
        double F2 = p.ro0_n * (1. + p.beta_n * (P + P_k - p.P_atm)) * Sn - roS_n;
        double F1P = p.ro0_w * p.beta_w * (1 - Sn);
        double F2P = p.ro0_n * p.beta_n * Sn;
        double F1S = (-1) * p.ro0_w * (1 + p.beta_w * (P - p.P_atm));
        double F2S = p.ro0_n * (1 + p.beta_n * (P + P_k - p.P_atm + (Sn * PkS)));
        double det = F1P * F2S - F1S * F2P;
        double rdet = rcp_rn(det);  // 1 / det
        P = P - rdet * (F2S * F1 - F1S * F2);
        Sn = Sn - rdet * (F1P * F2 - F2P * F1);
    }
}

// Three-phase fixed-iteration 3x3 Newton (three-phase.cpp:286-339) with the
// CPU path's zero capillary pressures folded; dF[4] = ro0_n*S_n keeps the
// reference's Jacobian entry (three-phase.cpp:319).
CAPAZ_HD void newton3(const DevParams& p, double& P, double& Sw, double& Sn, double roS_w, double roS_n,
                      double roS_g) {
    for (int w = 1; w <= p.iterations; w++) {
        double Sg = 1. - Sw - Sn;
        double rw = p.ro0_w * (1. + p.beta_w * (P - p.P_atm));  // dF[1]
        double rn = p.ro0_n * (1. + p.beta_n * (P - p.P_atm));  // dF[5]
        double gP = exact_div(p.ro0_g * P, p.P_atm, p.r_P_atm);
        double F1 = rw * Sw - roS_w;
        double F2 = rn * Sn - roS_n;
        double F3 = gP * Sg - roS_g;
        double dF[9];
        dF[0] = p.c_wb * Sw;
        dF[1] = rw;
        dF[2] = 0;
        dF[3] = p.c_nb * Sn;
        dF[4] = p.ro0_n * Sn;
        dF[5] = rn;
        dF[6] = exact_div(p.ro0_g * Sg, p.P_atm, p.r_P_atm);
        dF[7] = -gP;
        dF[8] = -gP;
        reverse3<true>(dF);
        P = P - (dF[0] * F1 + dF[1] * F2 + dF[2] * F3);
        Sw = Sw - (dF[3] * F1 + dF[4] * F2 + dF[5] * F3);
        Sn = Sn - (dF[6] * F1 + dF[7] * F2 + dF[8] * F3);
    }
}

// ---- energy closures (energy.cpp:4-143) ---------------------------------------------------

constexpr double kRoR = 2000.;
constexpr double kLam0W = 0.553, kLam0N = 0.14, kLam0G = 0.0237, kLam0R = 1.;
constexpr double kLamAW = 3E-3, kLamAN = 1E-3, kLamAG = 0.82;
constexpr double kC0W = 4.194E3, kC0N = 1.7E3, kC0G = 1E3, kC0R = 0.8E3;
constexpr double kCW = 1.15, kCW2 = 0.015, kCN = 3.4, kCG = 0.119, kCR = 0.75;

CAPAZ_HD double c_w_T(double T) { return kC0W - kCW * (T - kT0) + kCW2 * (T - kT0) * (T - kT0); }
CAPAZ_HD double c_n_T(double T) { return kC0N + kCN * (T - kT0); }
CAPAZ_HD double c_g_T(double T) { return kC0G + kCG * (T - kT0); }
CAPAZ_HD double c_r_T(double T) { return kC0R + kCR * (T - kT0); }

CAPAZ_HD double lambda_eff(double m, double Sw, double Sn, double Sg, double T) {
    double lw = kLam0W * (1 - kLamAW * (T - kT0));
    double ln = kLam0N * (1 - kLamAN * (T - kT0));
    double lg = kLam0G * pow_cl(exact_div(T, kT0, 1. / kT0), kLamAG);
    return m * (Sw * lw + Sn * ln + Sg * lg) + (1. - m) * kLam0R;
}

CAPAZ_HD double H_w_T(const DevParams& p, double T) {
    return p.H0_w + (T - kT0) * (kC0W - (T - kT0) * (kCW * 0.5 + exact_div(kCW2 * (T - kT0), 3., 1. / 3.)));
}
CAPAZ_HD double H_n_T(const DevParams& p, double T) { return p.H0_n + (T - kT0) * (kC0N + half(kCN * (T - kT0))); }
CAPAZ_HD double H_g_T(const DevParams& p, double T) { return p.H0_g + (T - kT0) * (kC0G + half(kCG * (T - kT0))); }
CAPAZ_HD double H_r_T(const DevParams& p, double T) { return p.H0_r + (T - kT0) * (kC0R + half(kCR * (T - kT0))); }

// ---- the configurations' energy closure (no reference counterpart) -----------------------
// Selected by p.ecfg: a constant conductivity of the medium and rock heat
// capacity, mu0 * exp(-a (T - T_ref)), and the water enthalpy as the integral
// of c_w (energy.cpp:114's cubic term has the opposite sign).

CAPAZ_HD double mu_cfg(const DevParams& p, double mu0, double T) { return mu0 * exp(-p.mu_T_a * (T - p.mu_T_ref)); }

CAPAZ_HD double H_w_of(const DevParams& p, double T) {
    if (p.ecfg)
        return p.H0_w + (T - kT0) * (kC0W - (T - kT0) * (kCW * 0.5 - exact_div(kCW2 * (T - kT0), 3., 1. / 3.)));
    return H_w_T(p, T);
}
CAPAZ_HD double H_r_of(const DevParams& p, double T) { return p.ecfg ? p.H0_r + (T - kT0) * p.c_r : H_r_T(p, T); }
CAPAZ_HD double c_r_of(const DevParams& p, double T) { return p.ecfg ? p.c_r : c_r_T(T); }

// mult_matrix_vector (gauss.cpp:94-104).
template <int N>
CAPAZ_HD void mult_matrix_vector(double* r, const double* m, const double* v) {
#pragma unroll
    for (int i = 0; i < N; i++) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < N; j++) s += m[j + N * i] * v[j];
        r[i] = s;
    }
}

// reverse_matrix<N> when no pivot search is needed: the general routine's
// operations in its order on the columns that can still reach the result.
// At step k those are the left columns after k and the right (inverse)
// columns up to N + k: the left columns up to k are never read again, and a
// right column after N + k is still the identity's (zero outside row k's
// divided zero), so the skipped updates only subtract zeros.  Column N + k
// is still the identity's at step k: its row-k entry 1 / f is RN(1/f) = rf
// itself, and its other entries 0 - rf * f1 are -(rf * f1).  The result has
// the general routine's bits up to the sign of entries that are exactly 0.
// Returns false, leaving `a` untouched, if some pivot has |f| < 1e-8.
template <int N>
CAPAZ_HD bool reverse_nopivot(double* a) {
    constexpr int N1 = 2 * N;
    const double e = 1E-8;
    double m[N][N1];
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int j = 0; j < N; j++) m[i][j] = a[j + N * i];
#pragma unroll
        for (int j = N; j < N1; j++) m[i][j] = (j == N + i) ? 1. : 0.;
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
        const double f = m[k][k];
        if (fabs(f) < e) return false;
        const double rf = rcp_rn_pivot(f);  // RN(1/f), as the general routine's 1. / f
#pragma unroll
        for (int j = k + 1; j < N + k; j++) m[k][j] = exact_div(m[k][j], f, rf);
        m[k][N + k] = rf;
#pragma unroll
        for (int i = 0; i < N; i++) {
            if (i != k) {
                const double f1 = m[i][k];
#pragma unroll
                for (int j = k + 1; j < N + k; j++) m[i][j] -= (m[k][j] * f1);
                m[i][N + k] = -(rf * f1);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) a[j + N * i] = m[i][j + N];
    return true;
}

// Energy-model 5x5 Newton in (S_w, S_n, S_g, P_w, T) (energy.cpp:370-463);
// H_* are refreshed after every iteration as assign_H does.  NOPIVOT: every
// inverse by reverse_nopivot; returns false as soon as one needs the pivot
// search (the arguments are then partly updated: the caller keeps copies and
// reruns the cell with NOPIVOT false, which always returns true).
template <bool NOPIVOT>
CAPAZ_HD bool newton_energy_t(const DevParams& p, double& Sw, double& Sn, double& Sg, double& P, double& T, double& Hw,
                              double& Hn, double& Hg, double& Hr, double m, double roS_w, double roS_n, double roS_g,
                              double E_new) {
    constexpr int n = 5;
    for (int w = 1; w <= p.iterations; w++) {
        double rw = ro_w_T(p, P, T), rn = ro_n_T(p, P, T), rg = ro_g_T(p, P, T);
        double dwP = p.ro0_w * p.beta_w, dwT = (-1) * p.ro0_w * kAlfaW;
        double dnP = p.ro0_n * p.beta_n, dnT = (-1) * p.ro0_n * kAlfaN;
        double dgP = exact_div(p.ro0_g, p.P_atm, p.r_P_atm) * (kT0 / T);
        double dgT = p.ro0_g * exact_div(P, p.P_atm, p.r_P_atm) * (-1) * (kT0 / T) / T;
        double F[5], dF[25], corr[5];
        F[0] = Sg + Sw + Sn - 1.;
        F[1] = rw * Sw - roS_w;
        F[2] = rn * Sn - roS_n;
        F[3] = rg * Sg - roS_g;
        F[4] = m * (Sw * (rw * Hw - P) + Sn * (rn * Hn - P) + Sg * (rg * Hg - P)) + (1. - m) * (p.ro_r * Hr - P) - E_new;
        dF[0] = 1.; dF[1] = 1.; dF[2] = 1.; dF[3] = 0.; dF[4] = 0.;
        dF[0 + n] = rw; dF[1 + n] = 0.; dF[2 + n] = 0.; dF[3 + n] = dwP * Sw; dF[4 + n] = dwT * Sw;
        dF[0 + n * 2] = 0.; dF[1 + n * 2] = rn; dF[2 + n * 2] = 0.; dF[3 + n * 2] = dnP * Sn; dF[4 + n * 2] = dnT * Sn;
        dF[0 + n * 3] = 0.; dF[1 + n * 3] = 0.; dF[2 + n * 3] = rg; dF[3 + n * 3] = dgP * Sg; dF[4 + n * 3] = dgT * Sg;
        dF[0 + n * 4] = m * (rw * Hw - P);
        dF[1 + n * 4] = m * (rn * Hn - P);
        dF[2 + n * 4] = m * (rg * Hg - P);
        dF[3 + n * 4] = m * (Sw * (dwP * Hw - 1.) + Sn * (dnP * Hn - 1.) + Sg * (dgP * Hg - 1.)) + (1. - m) * (-1);
        dF[4 + n * 4] = m * (Sw * (dwT * Hw + rw * c_w_T(T)) + Sn * (dnT * Hn + rn * c_n_T(T)) +
                             Sg * (dgT * Hg + rg * c_g_T(T))) +
                        (1. - m) * p.ro_r * c_r_of(p, T);
        if constexpr (NOPIVOT) {
            if (!reverse_nopivot<5>(dF)) return false;
        } else {
            reverse_matrix<5>(dF);
        }
        mult_matrix_vector<5>(corr, dF, F);
        Sw = Sw - corr[0];
        Sn = Sn - corr[1];
        Sg = Sg - corr[2];
        P = P - corr[3];
        T = T - corr[4];
        Hw = H_w_of(p, T);
        Hn = H_n_T(p, T);
        Hg = H_g_T(p, T);
        Hr = H_r_of(p, T);
    }
    return true;
}

CAPAZ_HD void newton_energy(const DevParams& p, double& Sw, double& Sn, double& Sg, double& P, double& T, double& Hw,
                            double& Hn, double& Hg, double& Hr, double m, double roS_w, double roS_n, double roS_g,
                            double E_new) {
    newton_energy_t<false>(p, Sw, Sn, Sg, P, T, Hw, Hn, Hg, Hr, m, roS_w, roS_n, roS_g, E_new);
}

// The two-phase energy model's Newton (configuration closure): energy.cpp:370-463's
// system without the gas row and column, 4 x 4 in (S_w, S_n, P_w, T), the
// densities at P_w for both phases as energy.cpp:388-394 has them, H_*
// refreshed after every iteration.  NOPIVOT as newton_energy_t.
template <bool NOPIVOT>
CAPAZ_HD bool newton_energy2_t(const DevParams& p, double& Sw, double& Sn, double& P, double& T, double& Hw,
                               double& Hn, double& Hr, double m, double roS_w, double roS_n, double E_new) {
    constexpr int n = 4;
    for (int w = 1; w <= p.iterations; w++) {
        double rw = ro_w_T(p, P, T), rn = ro_n_T(p, P, T);
        double dwP = p.ro0_w * p.beta_w, dwT = (-1) * p.ro0_w * kAlfaW;
        double dnP = p.ro0_n * p.beta_n, dnT = (-1) * p.ro0_n * kAlfaN;
        double F[4], dF[16], corr[4];
        F[0] = Sw + Sn - 1.;
        F[1] = rw * Sw - roS_w;
        F[2] = rn * Sn - roS_n;
        F[3] = m * (Sw * (rw * Hw - P) + Sn * (rn * Hn - P)) + (1. - m) * (p.ro_r * Hr - P) - E_new;
        dF[0] = 1.; dF[1] = 1.; dF[2] = 0.; dF[3] = 0.;
        dF[0 + n] = rw; dF[1 + n] = 0.; dF[2 + n] = dwP * Sw; dF[3 + n] = dwT * Sw;
        dF[0 + n * 2] = 0.; dF[1 + n * 2] = rn; dF[2 + n * 2] = dnP * Sn; dF[3 + n * 2] = dnT * Sn;
        dF[0 + n * 3] = m * (rw * Hw - P);
        dF[1 + n * 3] = m * (rn * Hn - P);
        dF[2 + n * 3] = m * (Sw * (dwP * Hw - 1.) + Sn * (dnP * Hn - 1.)) + (1. - m) * (-1);
        dF[3 + n * 3] = m * (Sw * (dwT * Hw + rw * c_w_T(T)) + Sn * (dnT * Hn + rn * c_n_T(T))) +
                        (1. - m) * p.ro_r * c_r_of(p, T);
        if constexpr (NOPIVOT) {
            if (!reverse_nopivot<4>(dF)) return false;
        } else {
            reverse_matrix<4>(dF);
        }
        mult_matrix_vector<4>(corr, dF, F);
        Sw = Sw - corr[0];
        Sn = Sn - corr[1];
        P = P - corr[2];
        T = T - corr[3];
        Hw = H_w_of(p, T);
        Hn = H_n_T(p, T);
        Hr = H_r_of(p, T);
    }
    return true;
}

// ---- three-phase Border_P faces (three-phase.cpp:383-422) ---------------------------------

// Sets (P_w, P_n, P_g) of a boundary cell in row j from the inward
// neighbour's P_w.
CAPAZ_HD void border_p3(const DevParams& p, int j, double P1, double& Pw, double& Pn, double& Pg) {
    if ((j != 0) && (j != p.Ny - 1)) {
        Pw = P1;
        Pn = P1;
        Pg = P1;
    } else if (j == 0) {
        Pw = exact_div(P1 - p.bp_w_num, p.bp_w_top_den, p.r_bp_w_top_den);
        Pn = exact_div(P1 - p.bp_n_num, p.bp_n_top_den, p.r_bp_n_top_den);
        Pg = exact_div(P1, p.bp_g_top_den, p.r_bp_g_top_den);
    } else {
        Pw = exact_div(P1 + p.bp_w_num, p.bp_w_bot_den, p.r_bp_w_bot_den);
        Pn = exact_div(P1 + p.bp_n_num, p.bp_n_bot_den, p.r_bp_n_bot_den);
        Pg = exact_div(P1, p.bp_g_bot_den, p.r_bp_g_bot_den);
    }
}

// Border_T (energy.cpp:343-366) with the configurations' face temperatures:
// top and bottom fixed where set (> 0), else, like the side faces, insulated
// (the inward cell's T1).
CAPAZ_HD double border_T_cfg(const DevParams& p, int j, double T1) {
    if (j == 0 && p.T_top > 0) return p.T_top;
    if (j == p.Ny - 1 && p.T_bottom > 0) return p.T_bottom;
    return T1;
}

// Two-phase Border_S window test at the top face (two-phase.cpp:175-181).
CAPAZ_HD bool in_source_window(const DevParams& p, int i, int k) {
    return (i >= p.Nx / 2 - p.source) && (i <= p.Nx / 2 + p.source) && (k >= p.Nz / 2 - p.source) &&
           (k <= p.Nz / 2 + p.source);
}

}  // namespace capaz
