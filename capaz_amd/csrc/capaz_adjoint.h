// Discrete adjoint of the fused three-phase step (Corey closures and the
// reference's van Genuchten closures), per cell.
//
// The forward step maps the state x_t (P_w, S_w, S_n in every cell; P_n, P_g
// in the boundary cells, where they persist) and the parameters (K, m) to
// x_{t+1}:
//   interior l:  A_a(l) = roS_a(x_t(l)) - dtm(l) * div_a(l),
//                x_{t+1}(l) = Newton(x_t(l), A(l))     (4 fixed iterations)
//   boundary b:  x_{t+1}(b) = Border(x_{t+1}(inward(b)))
// with div_a(l) the upwind 7-point flux divergence of phase a over the
// coefficients (P_a, Xi_a, ro_a) of l and its six neighbours
// (cpu.cpp:619-731) and dtm = dt / m.  Given the adjoint lambda_{t+1} of
// x_{t+1}, the adjoint step returns lambda_t and adds the step's share of
// dJ/d lnK and dJ/dm.  It runs in two passes over the cells:
//
//  1. adj_newton_cell (interior cells): folds the boundary cells' adjoints
//     into their inward cell (the transpose of Border), recomputes A(l) and
//     the Newton iterates, and runs the iteration backwards.  For
//     x' = x - u, u = M(x)^-1 F(x, A) with M the reference's Newton matrix
//     (three-phase.cpp:307-322, including its dF[4] = ro0_n*S_n entry) and
//     J = dF/dx the true Jacobian:
//         w = M^-T lambda',   lambda_A += w,
//         lambda_j = lambda'_j - w . (J[:, j] - (dM/dx_j) u).
//     It writes mu_a(l) = -dtm(l) * lambda_A_a(l), the adjoint reaching x_t(l)
//     through the initial guess and roS, and the porosity gradient.
//  2. adj_coef_cell (every cell): gathers mu over the transposed stencil --
//     the cell's own divergence and the face terms of its interior
//     neighbours that read its coefficients -- and chains the coefficients'
//     adjoint to the state (P_a, S_w, S_n) and to ln K.
//
// The upwind switches are differentiated piecewise, with the mean of the
// one-sided slopes at a zero driving force.  Nothing here needs to match
// the forward step bit for bit; the adjoint is checked against central
// finite differences of the forward step.
#pragma once

#include "capaz_bodies.h"

namespace capaz {

// Adjoint pass arrays (device or host).
struct AdjIO {
    const double *Pw, *Sw, *Sn, *Pn, *Pg, *m, *K, *dtm;  // x_t and the parameters
    const double *lPw1, *lSw1, *lSn1, *lPn1, *lPg1;       // lambda_{t+1}
    double *mu_w, *mu_n, *mu_g;                           // pass 1 -> pass 2 (interior cells)
    double *dPw, *dSw, *dSn;                              // pass 1 -> pass 2 (interior cells)
    double *lPw, *lSw, *lSn, *lPn, *lPg;                  // lambda_t
    double *g_lnK, *g_m;                                  // accumulated gradients
};

// Corey k_r and dk_r/ds at an effective saturation; N2: the exponent is 2
// (no pow in the code).
template <bool N2 = false>
CAPAZ_HD void corey_d(const DevParams& p, double s, double& kr, double& dkr) {
    if (s <= 0.) {
        kr = 0.;
        dkr = 0.;
    } else if (s >= 1.) {
        kr = 1.;
        dkr = 0.;
    } else if (N2 || p.corey_n == 2.) {
        kr = s * s;
        dkr = 2. * s;
    } else {
        kr = pow_cl(s, p.corey_n);
        dkr = p.corey_n * pow_cl(s, p.corey_n - 1.);
    }
}

// The reference's van Genuchten k_r (three-phase.cpp:146-186, as vg_k_w,
// vg_k_n and vg_k_g evaluate them) and their partial derivatives in the
// effective saturations x = Swe, y = Sne (z = Sge = 1 - x - y), by the chain
// rule on each power (a = p.vg_a, b = p.vg_b, c = p.vg_c; the powers are the
// forward step's pow_cl, whose log of a base the compiler computes once):
//   k_w = x^(1/2) t^2,           t = 1 - (1 - x^a)^b        (x >= 1e-3)
//   k_g = z^(1/2) v^c,           v = 1 - (1 - z)^a          (z >= 1e-3)
//   k_n = y k_nw k_ng / D,       D = (1 - x)(x + y)         (y >= 1e-3)
//         k_nw = (1 - x)^(1/2) (1 - x^a)^c,  k_ng = y^(1/2) s^2,  s = 1 - (1 - y^a)^b.
CAPAZ_HD void vg_d(const DevParams& p, double x, double y, double& kw, double& kn, double& kg, double& dkw_dx,
                   double& dkn_dx, double& dkn_dy, double& dkg_dz) {
    const double a = p.vg_a, b = p.vg_b, c = p.vg_c;
    const double z = 1. - x - y;
    kw = kn = kg = dkw_dx = dkn_dx = dkn_dy = dkg_dz = 0.;
    if (x >= 1e-3) {
        const double u = 1. - pow_cl(x, a), t = 1. - pow_cl(u, b);
        const double dt = b * pow_cl(u, b - 1.) * a * pow_cl(x, a - 1.);
        const double sx = sqrt(x);
        kw = sx * t * t;
        dkw_dx = 0.5 / sx * t * t + 2. * sx * t * dt;
    }
    if (z >= 1e-3) {
        const double v = 1. - pow_cl(1. - z, a), vc = pow_cl(v, c);
        const double dv = a * pow_cl(1. - z, a - 1.);
        const double sz = sqrt(z);
        kg = sz * vc;
        dkg_dz = 0.5 / sz * vc + sz * c * pow_cl(v, c - 1.) * dv;
    }
    if (y >= 1e-3) {
        const double w = 1. - pow_cl(x, a), sw = sqrt(1. - x);
        const double knw = sw * pow_cl(w, c);
        const double dknw = -0.5 / sw * pow_cl(w, c) - sw * c * pow_cl(w, c - 1.) * a * pow_cl(x, a - 1.);
        const double q = 1. - pow_cl(y, a), s = 1. - pow_cl(q, b);
        const double ds = b * pow_cl(q, b - 1.) * a * pow_cl(y, a - 1.);
        const double sy = sqrt(y);
        const double kng = sy * s * s;
        const double dkng = 0.5 / sy * s * s + 2. * sy * s * ds;
        const double D = (1. - x) * (x + y);
        kn = y * knw * kng / D;
        dkn_dx = y * kng * (dknw * D - knw * (1. - 2. * x - y)) / (D * D);
        dkn_dy = knw * ((kng + y * dkng) * D - y * kng * (1. - x)) / (D * D);
    }
}

// k_r of the model's closure (Corey or the reference's van Genuchten) and
// their partial derivatives: dkw = dk_w/dSwe, dkn_w / dkn_n = dk_n/dSwe /
// dSne, dkg = dk_g/dSge.  CL: 1 Corey, 0 van Genuchten, -1 p.closure at run
// time, 2 Corey with the exponent 2 (a kernel that knows its closure leaves
// the others' code out: the van Genuchten registers and the device pow).
template <int CL = -1>
CAPAZ_HD void kr3_d(const DevParams& p, double Swe, double Sne, double& kw, double& kn, double& kg, double& dkw,
                    double& dkn_w, double& dkn_n, double& dkg) {
    if (CL == 0 || (CL < 0 && p.closure == 0)) {
        vg_d(p, Swe, Sne, kw, kn, kg, dkw, dkn_w, dkn_n, dkg);
        return;
    }
    corey_d<CL == 2>(p, Swe, kw, dkw);
    corey_d<CL == 2>(p, Sne, kn, dkn_n);
    corey_d<CL == 2>(p, 1. - Swe - Sne, kg, dkg);
    dkn_w = 0.;
}

// A cell's coefficients and their derivatives with respect to its own state.
struct CoefD {
    double P[3], Xi[3], ro[3];
    double dXi_dSw[3], dXi_dSn[3];
    double dro_dP[3];
    bool inner;  // P_n = P_g = P_w (three-phase.cpp:225-235)
};

template <int CL = -1, class IO>
CAPAZ_HD CoefD coef_d(const DevParams& p, const IO& a, int64_t l, bool inner) {
    CoefD c;
    const double Pw = a.Pw[l], Sw = a.Sw[l], Sn = a.Sn[l], K = a.K[l];
    c.inner = inner;
    c.P[0] = Pw;
    c.P[1] = inner ? Pw : a.Pn[l];
    c.P[2] = inner ? Pw : a.Pg[l];
    const double Swe = (Sw - p.S_wr) * p.r_den3, Sne = (Sn - p.S_nr) * p.r_den3;
    double kw, dkw, kn, dkn_w, dkn_n, kg, dkg;
    kr3_d<CL>(p, Swe, Sne, kw, kn, kg, dkw, dkn_w, dkn_n, dkg);
    c.Xi[0] = -K * kw * p.r_mu_w;
    c.Xi[1] = -K * kn * p.r_mu_n;
    c.Xi[2] = -K * kg * p.r_mu_g;
    c.dXi_dSw[0] = -K * dkw * p.r_mu_w * p.r_den3;
    c.dXi_dSn[0] = 0.;
    c.dXi_dSw[1] = -K * dkn_w * p.r_mu_n * p.r_den3;
    c.dXi_dSn[1] = -K * dkn_n * p.r_mu_n * p.r_den3;
    c.dXi_dSw[2] = K * dkg * p.r_mu_g * p.r_den3;
    c.dXi_dSn[2] = c.dXi_dSw[2];
    c.ro[0] = p.ro0_w * (1. + p.beta_w * (c.P[0] - p.P_atm));
    c.ro[1] = p.ro0_n * (1. + p.beta_n * (c.P[1] - p.P_atm));
    c.ro[2] = p.ro0_g * c.P[2] * p.r_P_atm;
    c.dro_dP[0] = p.c_wb;
    c.dro_dP[1] = p.c_nb;
    c.dro_dP[2] = p.ro0_g * p.r_P_atm;
    return c;
}

// Values only (neighbours in the gathers).
template <int CL = -1, class IO>
CAPAZ_HD Coef coef_v(const DevParams& p, const IO& a, int64_t l, bool inner) {
    Coef c;
    const double Pw = a.Pw[l], Sw = a.Sw[l], Sn = a.Sn[l], K = a.K[l];
    c.P[0] = Pw;
    c.P[1] = inner ? Pw : a.Pn[l];
    c.P[2] = inner ? Pw : a.Pg[l];
    const double Swe = (Sw - p.S_wr) * p.r_den3, Sne = (Sn - p.S_nr) * p.r_den3;
    double kw, kn, kg, d0, d1, d2, d3;
    kr3_d<CL>(p, Swe, Sne, kw, kn, kg, d0, d1, d2, d3);
    c.Xi[0] = -K * kw * p.r_mu_w;
    c.Xi[1] = -K * kn * p.r_mu_n;
    c.Xi[2] = -K * kg * p.r_mu_g;
    c.ro[0] = p.ro0_w * (1. + p.beta_w * (c.P[0] - p.P_atm));
    c.ro[1] = p.ro0_n * (1. + p.beta_n * (c.P[1] - p.P_atm));
    c.ro[2] = p.ro0_g * c.P[2] * p.r_P_atm;
    return c;
}

// 3x3 solve M^T w = r by cofactors (M row-major).
CAPAZ_HD void solve3_transposed(const double* M, const double* r, double* w) {
    // M^T = [[M0, M3, M6], [M1, M4, M7], [M2, M5, M8]]
    const double a = M[0], b = M[3], c = M[6];
    const double d = M[1], e = M[4], f = M[7];
    const double g = M[2], h = M[5], i = M[8];
    const double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    const double rd = 1. / det;
    const double D = -(b * i - c * h), E = a * i - c * g, F = -(a * h - b * g);
    const double G = b * f - c * e, H = -(a * f - c * d), I = a * e - b * d;
    // inverse of T = [[a,b,c],[d,e,f],[g,h,i]] is adj / det with adj = [[A, D, G], [B, E, H], [C, F, I]]
    w[0] = (A * r[0] + D * r[1] + G * r[2]) * rd;
    w[1] = (B * r[0] + E * r[1] + H * r[2]) * rd;
    w[2] = (C * r[0] + F * r[1] + I * r[2]) * rd;
}

// The reference Newton matrix M (rows F1..F3, columns P, S_w, S_n) at x.
CAPAZ_HD void newton3_matrix(const DevParams& p, double P, double Sw, double Sn, double* M) {
    const double Sg = 1. - Sw - Sn;
    const double gP = p.ro0_g * P * p.r_P_atm;
    M[0] = p.c_wb * Sw;
    M[1] = p.ro0_w * (1. + p.beta_w * (P - p.P_atm));
    M[2] = 0.;
    M[3] = p.c_nb * Sn;
    M[4] = p.ro0_n * Sn;
    M[5] = p.ro0_n * (1. + p.beta_n * (P - p.P_atm));
    M[6] = p.ro0_g * Sg * p.r_P_atm;
    M[7] = -gP;
    M[8] = -gP;
}

// Adjoint of the fixed-iteration Newton (newton3) from (P0, Sw0, Sn0) with
// right-hand sides A: given lam (adjoint of the result, overwritten with the
// adjoint of the initial guess), adds the adjoint of A to lA.  The forward
// iterates are recomputed first (the same iteration as newton3, solved by
// cofactors).  IT > 0: that many iterations, unrolled so the iterates stay in
// registers; IT == 0: p.iterations (at most kMaxAdjIter) in a loop.
constexpr int kMaxAdjIter = 8;

template <int IT>
CAPAZ_HD void newton3_adjoint(const DevParams& p, double P0, double Sw0, double Sn0, const double* A, double* lam,
                              double* lA) {
    constexpr int NX = IT > 0 ? IT : kMaxAdjIter;
    double xs[NX + 1][3];
    const int it = IT > 0 ? IT : (p.iterations < kMaxAdjIter ? p.iterations : kMaxAdjIter);
    xs[0][0] = P0;
    xs[0][1] = Sw0;
    xs[0][2] = Sn0;
#pragma unroll
    for (int k = 0; k < NX; ++k) {
        if (k >= it) break;
        const double P = xs[k][0], Sw = xs[k][1], Sn = xs[k][2];
        double M[9];
        newton3_matrix(p, P, Sw, Sn, M);
        const double F[3] = {M[1] * Sw - A[0], M[5] * Sn - A[1], -M[7] * (1. - Sw - Sn) - A[2]};
        // u = M^-1 F, as the transposed solve of M^T
        double MT[9] = {M[0], M[3], M[6], M[1], M[4], M[7], M[2], M[5], M[8]};
        double u[3];
        solve3_transposed(MT, F, u);
        xs[k + 1][0] = P - u[0];
        xs[k + 1][1] = Sw - u[1];
        xs[k + 1][2] = Sn - u[2];
    }
    const double rg = p.ro0_g * p.r_P_atm;
#pragma unroll
    for (int kk = NX - 1; kk >= 0; --kk) {
        if (kk >= it) continue;
        const int k = kk;
        const double P = xs[k][0], Sw = xs[k][1], Sn = xs[k][2];
        const double u0 = P - xs[k + 1][0], u1 = Sw - xs[k + 1][1], u2 = Sn - xs[k + 1][2];
        double M[9];
        newton3_matrix(p, P, Sw, Sn, M);
        double w[3];
        solve3_transposed(M, lam, w);
        lA[0] += w[0];
        lA[1] += w[1];
        lA[2] += w[2];
        // columns of J - (dM/dx_j) u, rows F1..F3
        // j = P:   J = (c_wb Sw, c_nb Sn, rg Sg);  (dM/dP) u = (c_wb u1, c_nb u2, -rg (u1 + u2))
        // j = S_w: J = (rw, 0, -gP);              (dM/dSw) u = (c_wb u0, 0, -rg u0)
        // j = S_n: J = (0, rn, -gP);              (dM/dSn) u = (0, c_nb u0 + ro0_n u1, -rg u0)
        const double gP = rg * P;
        const double cP0 = M[0] - p.c_wb * u1, cP1 = M[3] - p.c_nb * u2, cP2 = M[6] + rg * (u1 + u2);
        const double cW0 = M[1] - p.c_wb * u0, cW2 = -gP + rg * u0;
        const double cN1 = M[5] - (p.c_nb * u0 + p.ro0_n * u1), cN2 = -gP + rg * u0;
        lam[0] -= w[0] * cP0 + w[1] * cP1 + w[2] * cP2;
        lam[1] -= w[0] * cW0 + w[2] * cW2;
        lam[2] -= w[1] * cN1 + w[2] * cN2;
    }
}

// Boundary cells whose inward neighbour is the interior cell (i, j, k), as
// boundary_epilogue enumerates them; f(bi, bj, bk) for each.
template <class Fn>
CAPAZ_HD void for_boundary_of(const DevParams& p, int i, int j, int k, Fn&& f) {
    int ox[3] = {0, 0, 0}, oy[3] = {0, 0, 0}, oz[3] = {0, 0, 0};
    int nx = 1, ny = 1, nz = 1;
    if (p.Nx > 2 && i == 1) ox[nx++] = -1;
    if (p.Nx > 2 && i == p.Nx - 2) ox[nx++] = +1;
    if (j == 1) oy[ny++] = -1;
    if (j == p.Ny - 2) oy[ny++] = +1;
    if (p.Nz > 2 && k == 1) oz[nz++] = -1;
    if (p.Nz > 2 && k == p.Nz - 2) oz[nz++] = +1;
    for (int a = 0; a < nz; ++a)
        for (int b = 0; b < ny; ++b)
            for (int c = 0; c < nx; ++c) {
                if (a == 0 && b == 0 && c == 0) continue;
                f(i + ox[c], j + oy[b], k + oz[a]);
            }
}

// The transpose of Border for one boundary cell b (row bj) whose inward cell
// is l: folds b's adjoint lb = (P_w, S_w, S_n, P_n, P_g) into l's (P_w, S_w,
// S_n) adjoint lam (border_p3 and Border_S, three-phase.cpp:345-427).
CAPAZ_HD void border_fold(const DevParams& p, int bj, const double* lb, double* lam) {
    double dw = 1., dn = 1., dg = 1.;  // d(P_w, P_n, P_g)_b / dP_w(l)
    if (bj == 0) {
        dw = p.r_bp_w_top_den;
        dn = p.r_bp_n_top_den;
        dg = p.r_bp_g_top_den;
    } else if (bj == p.Ny - 1) {
        dw = p.r_bp_w_bot_den;
        dn = p.r_bp_n_bot_den;
        dg = p.r_bp_g_bot_den;
    }
    lam[0] += dw * lb[0] + dn * lb[3] + dg * lb[4];
    if (!(bj == 0 && p.source > 0)) {  // the injected saturations are constants
        lam[1] += lb[1];
        lam[2] += lb[2];
    }
}

// Pass 1's backward Newton at interior cell l with lam = the adjoint of
// x_{t+1}(l) including its boundary cells, once the step's densities ro,
// state (Pw, Sw, Sn) and flux divergence div of the cell are known (IT:
// newton3_adjoint's).  Writes mu, the direct adjoint d and the porosity
// gradient of the cell.
template <int IT>
CAPAZ_HD void adj_newton_core(const DevParams& p, const AdjIO& a, int64_t l, const double* ro, double Pw, double Sw,
                              double Sn, const double* div, double* lam) {
    const double Sg = 1. - Sw - Sn, dtm = a.dtm[l];
    double A[3];
    A[0] = ro[0] * Sw - dtm * div[0];
    A[1] = ro[1] * Sn - dtm * div[1];
    A[2] = ro[2] * Sg - dtm * div[2];
    double lA[3] = {0., 0., 0.};
    newton3_adjoint<IT>(p, Pw, Sw, Sn, A, lam, lA);
    a.mu_w[l] = -dtm * lA[0];
    a.mu_n[l] = -dtm * lA[1];
    a.mu_g[l] = -dtm * lA[2];
    // dA/dm = div * dtm / m
    a.g_m[l] += (lA[0] * div[0] + lA[1] * div[1] + lA[2] * div[2]) * dtm / a.m[l];
    // roS_a(x_t(l)) and the Newton initial guess
    const double rg = p.ro0_g * p.r_P_atm;
    a.dPw[l] = lam[0] + lA[0] * p.c_wb * Sw + lA[1] * p.c_nb * Sn + lA[2] * rg * Sg;
    a.dSw[l] = lam[1] + lA[0] * ro[0] - lA[2] * ro[2];
    a.dSn[l] = lam[2] + lA[1] * ro[1] - lA[2] * ro[2];
}

// Pass 1 at interior cell (i, j, k), l its index, reading lambda_{t+1} from
// the arrays.
template <int IT>
CAPAZ_HD void adj_newton_tail(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l, const double* ro,
                              double Pw, double Sw, double Sn, const double* div) {
    double lam[3] = {a.lPw1[l], a.lSw1[l], a.lSn1[l]};
    for_boundary_of(p, i, j, k, [&](int bi, int bj, int bk) {
        const int64_t lb = (int64_t)bi + (int64_t)bj * p.sy + (int64_t)bk * p.sz;
        const double v[5] = {a.lPw1[lb], a.lSw1[lb], a.lSn1[lb], a.lPn1[lb], a.lPg1[lb]};
        border_fold(p, bj, v, lam);
    });
    adj_newton_core<IT>(p, a, l, ro, Pw, Sw, Sn, div, lam);
}

// Pass 1 at interior cell (i, j, k) from the flux divergence recorded by the
// forward step (div_a at div + a*stride): the densities are recomputed with
// the step's own expressions, so A is the step's A to the bit.
template <int IT>
CAPAZ_HD void adj_newton_recorded(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l,
                                  const double* div_rec, int64_t stride) {
    const double Pw = a.Pw[l];
    const double ro[3] = {ro_w_of(p, Pw), ro_n_of(p, Pw), ro_g_of(p, Pw)};
    const double div[3] = {div_rec[l], div_rec[stride + l], div_rec[2 * stride + l]};
    adj_newton_tail<IT>(p, a, i, j, k, l, ro, Pw, a.Sw[l], a.Sn[l], div);
}

// Pass 1 at interior cell (i, j, k), reading its neighbours from memory.
template <int IT>
CAPAZ_HD void adj_newton_cell(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l) {
    const Coef c = coef_v(p, a, l, true);
    auto nb = [&](int ii, int jj, int kk) {
        const int64_t q = (int64_t)ii + (int64_t)jj * p.sy + (int64_t)kk * p.sz;
        return coef_v(p, a, q, interior3(p, ii, jj, kk));
    };
    Coef xm = c, xp = c, zm = c, zp = c;
    if (!(p.Nx < 2)) {
        xm = nb(i - 1, j, k);
        xp = nb(i + 1, j, k);
    }
    if (!(p.Nz < 2)) {
        zm = nb(i, j, k - 1);
        zp = nb(i, j, k + 1);
    }
    const Coef ym = nb(i, j - 1, k), yp = nb(i, j + 1, k);
    double div[3];
    for (int ph = 0; ph < 3; ++ph) div[ph] = coef_divergence(p, ph, c, xm, xp, ym, yp, zm, zp);
    adj_newton_tail<IT>(p, a, i, j, k, l, c.ro, a.Pw[l], a.Sw[l], a.Sn[l], div);
}

// Face-term derivatives of directed() for one cell's flux along one axis.
// x1, x2: the cell's minus / plus driving forces; L = -Xi*ro of the cell (c)
// and of its minus (m) / plus (p) neighbours; rh: 1 / the axis spacing.
// At a zero driving force the upwind choice switches: there the derivative
// of max(x, 0) is taken as 1/2 (the mean of the two one-sided slopes, which
// is what a central difference measures; states that are uniform along an
// axis put many faces exactly there).
CAPAZ_HD double upwind_w(double x) { return x > 0. ? 1. : (x < 0. ? 0. : 0.5); }

struct DirD {
    double dx1, dx2, dLc, dLm, dLp;
};
CAPAZ_HD DirD directed_d(double x1, double x2, double Lc, double Lm, double Lp, double rh) {
    DirD d;
    const double w1 = upwind_w(x1), w2 = upwind_w(x2);
    d.dx1 = (-(1. - w1) * Lc - w1 * Lm) * rh;
    d.dx2 = (w2 * Lc + (1. - w2) * Lp) * rh;
    d.dLc = ((x2 > 0. ? x2 : 0.) - (x1 < 0. ? x1 : 0.)) * rh;
    d.dLm = -(x1 > 0. ? x1 : 0.) * rh;
    d.dLp = (x2 < 0. ? x2 : 0.) * rh;
    return d;
}

// A neighbour along one axis as pass 2 reads it: its coefficients, its mu
// (zero unless it is an interior cell) and whether it exists.
struct AdjNb {
    Coef c;
    double mu[3];
    bool exists, interior;
};

// Pass 2, one axis of cell c: adds to lP, lL (L = -Xi ro) and lro the
// derivative of the cell's own flux (weighted by its mu, if interior) and of
// its interior neighbours' fluxes that read the cell.
CAPAZ_HD void adj_coef_axis(const DevParams& p, const CoefD& c, const double* mu_self, const AdjNb& m,
                            const AdjNb& q, double rh, bool grav, double* lP, double* lL, double* lro) {
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int ph = 0; ph < 3; ++ph) {
        const double Lc = -c.Xi[ph] * c.ro[ph];
        if (c.inner) {  // the cell's own flux along this axis (both neighbours exist)
            const double gr = grav ? p.g * c.ro[ph] : 0.;
            const double x1 = -((c.P[ph] - m.c.P[ph]) * rh) + gr, x2 = -((q.c.P[ph] - c.P[ph]) * rh) + gr;
            const DirD d = directed_d(x1, x2, Lc, -m.c.Xi[ph] * m.c.ro[ph], -q.c.Xi[ph] * q.c.ro[ph], rh);
            lP[ph] += mu_self[ph] * (d.dx2 - d.dx1) * rh;  // dx1/dP_c = -1/h, dx2/dP_c = +1/h
            lL[ph] += mu_self[ph] * d.dLc;
            if (grav) lro[ph] += mu_self[ph] * (d.dx1 + d.dx2) * p.g;
        }
        if (q.exists && q.interior) {  // plus neighbour's flux: this cell is its minus neighbour
            const double gr = grav ? p.g * q.c.ro[ph] : 0.;
            const double x1 = -((q.c.P[ph] - c.P[ph]) * rh) + gr;
            const double Lq = -q.c.Xi[ph] * q.c.ro[ph];
            const double w1 = upwind_w(x1);
            const double dx1 = (-(1. - w1) * Lq - w1 * Lc) * rh;
            lP[ph] += q.mu[ph] * dx1 * rh;  // dx1/dP_m = +1/h
            lL[ph] += q.mu[ph] * (-(x1 > 0. ? x1 : 0.) * rh);
        }
        if (m.exists && m.interior) {  // minus neighbour's flux: this cell is its plus neighbour
            const double gr = grav ? p.g * m.c.ro[ph] : 0.;
            const double x2 = -((c.P[ph] - m.c.P[ph]) * rh) + gr;
            const double Lq = -m.c.Xi[ph] * m.c.ro[ph];
            const double w2 = upwind_w(x2);
            const double dx2 = (w2 * Lq + (1. - w2) * Lc) * rh;
            lP[ph] += m.mu[ph] * (-dx2 * rh);  // dx2/dP_p = -1/h
            lL[ph] += m.mu[ph] * ((x2 < 0. ? x2 : 0.) * rh);
        }
    }
}

// Pass 2, after the axes: the coefficients' adjoint chained to the state and
// ln K of cell l, plus pass 1's share if the cell is interior: out =
// lambda_t(l) (P_w, S_w, S_n, P_n, P_g); returns the cell's ln K gradient.
CAPAZ_HD double adj_coef_lambda(const AdjIO& a, int64_t l, const CoefD& c, const double* lP, const double* lL,
                                const double* lro, double* out) {
    double lPw = 0., lSw = 0., lSn = 0., lPn = 0., lPg = 0., glnK = 0.;
    for (int ph = 0; ph < 3; ++ph) {
        const double lXi = -lL[ph] * c.ro[ph];
        const double lr = lro[ph] - lL[ph] * c.Xi[ph];
        const double lp = lP[ph] + lr * c.dro_dP[ph];
        if (ph == 0 || c.inner)
            lPw += lp;
        else if (ph == 1)
            lPn += lp;
        else
            lPg += lp;
        lSw += lXi * c.dXi_dSw[ph];
        lSn += lXi * c.dXi_dSn[ph];
        glnK += lXi * c.Xi[ph];  // dXi/dlnK = Xi
    }
    if (c.inner) {
        lPw += a.dPw[l];
        lSw += a.dSw[l];
        lSn += a.dSn[l];
    }
    out[0] = lPw;
    out[1] = lSw;
    out[2] = lSn;
    out[3] = lPn;
    out[4] = lPg;
    return glnK;
}

// ... stored into the lambda_t arrays and the ln K gradient.
CAPAZ_HD void adj_coef_finish(const DevParams& p, const AdjIO& a, int64_t l, const CoefD& c, const double* lP,
                              const double* lL, const double* lro) {
    (void)p;
    double v[5];
    const double glnK = adj_coef_lambda(a, l, c, lP, lL, lro, v);
    a.lPw[l] = v[0];
    a.lSw[l] = v[1];
    a.lSn[l] = v[2];
    a.lPn[l] = v[3];
    a.lPg[l] = v[4];
    a.g_lnK[l] += glnK;
}

// Pass 2 at cell (i, j, k), l its index, reading its neighbours from memory
// (CL: as kr3_d).
template <int CL = -1>
CAPAZ_HD AdjNb adj_nb_at(const DevParams& p, const AdjIO& a, int i, int j, int k) {
    AdjNb n;
    n.exists = i >= 0 && i < p.Nx && j >= 0 && j < p.Ny && k >= 0 && k < p.Nz;
    n.interior = false;
    n.mu[0] = n.mu[1] = n.mu[2] = 0.;
    n.c = Coef{};
    if (!n.exists) return n;
    const int64_t q = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz;
    n.interior = interior3(p, i, j, k);
    n.c = coef_v<CL>(p, a, q, n.interior);
    if (n.interior) {
        n.mu[0] = a.mu_w[q];
        n.mu[1] = a.mu_n[q];
        n.mu[2] = a.mu_g[q];
    }
    return n;
}

template <int CL = -1>
CAPAZ_HD void adj_coef_cell(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l) {
    const bool inner = interior3(p, i, j, k);
    const CoefD c = coef_d<CL>(p, a, l, inner);
    const double mu_self[3] = {inner ? a.mu_w[l] : 0., inner ? a.mu_n[l] : 0., inner ? a.mu_g[l] : 0.};
    double lP[3] = {0., 0., 0.}, lL[3] = {0., 0., 0.}, lro[3] = {0., 0., 0.};
    if (!(p.Nx < 2))
        adj_coef_axis(p, c, mu_self, adj_nb_at<CL>(p, a, i - 1, j, k), adj_nb_at<CL>(p, a, i + 1, j, k), p.r_hx, false, lP,
                      lL, lro);
    adj_coef_axis(p, c, mu_self, adj_nb_at<CL>(p, a, i, j - 1, k), adj_nb_at<CL>(p, a, i, j + 1, k), p.r_hy, true, lP, lL,
                  lro);
    if (!(p.Nz < 2))
        adj_coef_axis(p, c, mu_self, adj_nb_at<CL>(p, a, i, j, k - 1), adj_nb_at<CL>(p, a, i, j, k + 1), p.r_hz, false, lP,
                      lL, lro);
    adj_coef_finish(p, a, l, c, lP, lL, lro);
}

// ---- two phases (two-phase.cpp with the Corey closure) ------------------------
//
// The state is (P_w, S_n) per cell; with no capillary pressure P_n = P_w in
// every cell, so a cell's coefficients depend on its P_w (densities) and S_n
// (k_r).  The step's Newton (newton2, two-phase.cpp:25-56) uses the true
// Jacobian M = dF/dx of F = (ro_w(P)(1 - S_n) - A_w, ro_n(P) S_n - A_n), so
// its backward iteration is the three-phase one with J = M.  Border
// (two-phase.cpp:163-215): S_n from the inward cell (the injected value in
// the top face's source window), P_w from the inward cell on the sides,
// P_atm on the top, and on the bottom the fixed P_bottom or the inward cell's
// new P_w plus (ro_n S_n + ro_w (1 - S_n)) m g h_y with this step's
// densities of the inward cell's old state and its new S_n.
// Field use in AdjIO: Pw, Sn (state), lPw1/lSn1, mu_w/mu_n, dPw/dSn, lPw/lSn.

struct CoefD2 {
    double P, Xi[2], ro[2];
    double dXi_dSn[2], dro_dP[2];
};

template <bool N2 = false, class IO>
CAPAZ_HD CoefD2 coef_d2(const DevParams& p, const IO& a, int64_t l) {
    CoefD2 c;
    const double Pw = a.Pw[l], Sn = a.Sn[l], K = a.K[l];
    c.P = Pw;
    const double se = (1. - Sn - p.S_wr) * p.r_den2c;  // dse/dS_n = -1/den2c
    double kw, dkw, kn, dkn;
    corey_d<N2>(p, se, kw, dkw);
    corey_d<N2>(p, 1. - se, kn, dkn);
    c.Xi[0] = -K * kw * p.r_mu_w;
    c.Xi[1] = -K * kn * p.r_mu_n;
    c.dXi_dSn[0] = K * dkw * p.r_mu_w * p.r_den2c;
    c.dXi_dSn[1] = -K * dkn * p.r_mu_n * p.r_den2c;
    c.ro[0] = p.ro0_w * (1. + p.beta_w * (Pw - p.P_atm));
    c.ro[1] = p.ro0_n * (1. + p.beta_n * (Pw - p.P_atm));
    c.dro_dP[0] = p.c_wb;
    c.dro_dP[1] = p.c_nb;
    return c;
}

// newton2's matrix (rows F1, F2; columns P, S_n) at (P, S_n).
CAPAZ_HD void newton2_matrix(const DevParams& p, double P, double Sn, double* M) {
    M[0] = p.c_wb * (1. - Sn);
    M[1] = -p.ro0_w * (1. + p.beta_w * (P - p.P_atm));
    M[2] = p.c_nb * Sn;
    M[3] = p.ro0_n * (1. + p.beta_n * (P - p.P_atm));
}

// The forward iterates of newton2 from x0 = (P, S_n) with right-hand sides
// A (xs[0..it]), solved by Cramer's rule.
template <int IT>
CAPAZ_HD int newton2_iterates(const DevParams& p, double P0, double Sn0, const double* A,
                              double (&xs)[(IT > 0 ? IT : kMaxAdjIter) + 1][2]) {
    constexpr int NX = IT > 0 ? IT : kMaxAdjIter;
    const int it = IT > 0 ? IT : (p.iterations < kMaxAdjIter ? p.iterations : kMaxAdjIter);
    xs[0][0] = P0;
    xs[0][1] = Sn0;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int k = 0; k < NX; ++k) {
        if (k >= it) break;
        const double P = xs[k][0], Sn = xs[k][1];
        double M[4];
        newton2_matrix(p, P, Sn, M);
        const double F0 = -M[1] * (1. - Sn) - A[0], F1 = M[3] * Sn - A[1];
        const double rd = 1. / (M[0] * M[3] - M[1] * M[2]);
        xs[k + 1][0] = P - (M[3] * F0 - M[1] * F1) * rd;
        xs[k + 1][1] = Sn - (M[0] * F1 - M[2] * F0) * rd;
    }
    return it;
}

// Backward newton2 over the iterates xs: lam (the adjoint of xs[it]) becomes
// the adjoint of xs[0]; the adjoint of A is added to lA.  Per iteration,
// w = M^-T lam, lA += w, lam_j -= w . (J[:, j] - (dM/dx_j) u) with J = M:
//   j = P:   J = (c_wb (1 - S_n), c_nb S_n), (dM/dP) u = (-c_wb u1, c_nb u1)
//   j = S_n: J = (-ro_w, ro_n),              (dM/dS_n) u = (-c_wb u0, c_nb u0)
template <int IT>
CAPAZ_HD void newton2_backward(const DevParams& p, const double (&xs)[(IT > 0 ? IT : kMaxAdjIter) + 1][2], int it,
                               double* lam, double* lA) {
    constexpr int NX = IT > 0 ? IT : kMaxAdjIter;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int kk = NX - 1; kk >= 0; --kk) {
        if (kk >= it) continue;
        const double P = xs[kk][0], Sn = xs[kk][1];
        const double u0 = P - xs[kk + 1][0], u1 = Sn - xs[kk + 1][1];
        double M[4];
        newton2_matrix(p, P, Sn, M);
        // M^T w = lam: [[M0, M2], [M1, M3]] w = lam
        const double rd = 1. / (M[0] * M[3] - M[1] * M[2]);
        const double w0 = (M[3] * lam[0] - M[2] * lam[1]) * rd;
        const double w1 = (M[0] * lam[1] - M[1] * lam[0]) * rd;
        lA[0] += w0;
        lA[1] += w1;
        lam[0] -= w0 * (M[0] + p.c_wb * u1) + w1 * (M[2] - p.c_nb * u1);
        lam[1] -= w0 * (M[1] + p.c_wb * u0) + w1 * (M[3] - p.c_nb * u0);
    }
}

// Pass 1 of the two-phase adjoint at interior cell (i, j, k), l its index:
// the step's densities ro and flux divergence div of the cell are known.
// Folds the boundary cells' adjoints into the cell's new state (the transpose
// of Border), runs newton2 backwards and writes mu, the direct adjoint d and
// the porosity gradient.
template <int IT>
CAPAZ_HD void adj2_newton_core(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l, const double* ro,
                               const double* div) {
    const double Pw = a.Pw[l], Sn = a.Sn[l], dtm = a.dtm[l], m = a.m[l];
    double A[2];
    A[0] = ro[0] * (1. - Sn) - dtm * div[0];
    A[1] = ro[1] * Sn - dtm * div[1];
    double xs[(IT > 0 ? IT : kMaxAdjIter) + 1][2];
    const int it = newton2_iterates<IT>(p, Pw, Sn, A, xs);
    const double Sn1 = xs[it][1];
    double lam[2] = {a.lPw1[l], a.lSn1[l]};
    double dP_extra = 0., gm_extra = 0.;
    const double ghy = p.g * p.hy;
    for_boundary_of(p, i, j, k, [&](int bi, int bj, int bk) {
        const int64_t lb = (int64_t)bi + (int64_t)bj * p.sy + (int64_t)bk * p.sz;
        const double lbP = a.lPw1[lb], lbS = a.lSn1[lb];
        if (!(bj == 0 && p.source > 0 && in_source_window(p, bi, bk))) lam[1] += lbS;
        if (bj != 0 && bj != p.Ny - 1) {
            lam[0] += lbP;
        } else if (bj == p.Ny - 1 && !(p.P_bottom > 0)) {
            // P_b = P_w(new) + (ro_n S_n(new) + ro_w (1 - S_n(new))) m g h_y
            lam[0] += lbP;
            lam[1] += lbP * (ro[1] - ro[0]) * m * ghy;
            dP_extra += lbP * (p.c_nb * Sn1 + p.c_wb * (1. - Sn1)) * m * ghy;
            gm_extra += lbP * (ro[1] * Sn1 + ro[0] * (1. - Sn1)) * ghy;
        }
    });
    double lA[2] = {0., 0.};
    newton2_backward<IT>(p, xs, it, lam, lA);
    a.mu_w[l] = -dtm * lA[0];
    a.mu_n[l] = -dtm * lA[1];
    a.g_m[l] += (lA[0] * div[0] + lA[1] * div[1]) * dtm / m + gm_extra;
    a.dPw[l] = lam[0] + lA[0] * p.c_wb * (1. - Sn) + lA[1] * p.c_nb * Sn + dP_extra;
    a.dSn[l] = lam[1] - lA[0] * ro[0] + lA[1] * ro[1];
}

// Pass 1 from the flux divergence recorded by the forward step.
template <int IT>
CAPAZ_HD void adj2_newton_recorded(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l,
                                   const double* div_rec, int64_t stride) {
    const double Pw = a.Pw[l];
    const double ro[2] = {ro_w_of(p, Pw), ro_n_of(p, Pw)};
    const double div[2] = {div_rec[l], div_rec[stride + l]};
    adj2_newton_core<IT>(p, a, i, j, k, l, ro, div);
}

// Pass 1 recomputing the cell's divergence with the step's own arithmetic
// (the host build's path).
template <int IT>
CAPAZ_HD void adj2_newton_cell(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l) {
    StepIO io{};
    io.Pw = a.Pw;
    io.Sn = a.Sn;
    io.m = a.m;
    io.K = a.K;
    const Coef c = cell_coef(p, io, l, true);
    auto nb = [&](int ii, int jj, int kk) {
        const int64_t q = (int64_t)ii + (int64_t)jj * p.sy + (int64_t)kk * p.sz;
        return cell_coef(p, io, q, cell_inner(p, ii, jj, kk));
    };
    Coef xm = c, xp = c, zm = c, zp = c;
    if (!(p.Nx < 2)) {
        xm = nb(i - 1, j, k);
        xp = nb(i + 1, j, k);
    }
    if (!(p.Nz < 2)) {
        zm = nb(i, j, k - 1);
        zp = nb(i, j, k + 1);
    }
    const Coef ym = nb(i, j - 1, k), yp = nb(i, j + 1, k);
    const double div[2] = {coef_divergence(p, 0, c, xm, xp, ym, yp, zm, zp),
                           coef_divergence(p, 1, c, xm, xp, ym, yp, zm, zp)};
    adj2_newton_core<IT>(p, a, i, j, k, l, c.ro, div);
}

// Pass 2 of the two-phase adjoint at cell (i, j, k), l its index (every
// cell): the transposed stencil of the fluxes that read the cell's
// coefficients (its own, if interior, and its interior neighbours'), chained
// to (P_w, S_n) and ln K, plus pass 1's direct share if interior.
struct AdjNb2 {
    double P, Xi[2], ro[2], mu[2];
    bool exists, interior;
};

template <bool N2 = false>
CAPAZ_HD AdjNb2 adj2_nb_at(const DevParams& p, const AdjIO& a, int i, int j, int k) {
    AdjNb2 n{};
    n.exists = i >= 0 && i < p.Nx && j >= 0 && j < p.Ny && k >= 0 && k < p.Nz;
    if (!n.exists) return n;
    const int64_t q = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz;
    n.interior = interior_newton2(p, i, j, k);
    const CoefD2 c = coef_d2<N2>(p, a, q);
    n.P = c.P;
    n.Xi[0] = c.Xi[0];
    n.Xi[1] = c.Xi[1];
    n.ro[0] = c.ro[0];
    n.ro[1] = c.ro[1];
    if (n.interior) {
        n.mu[0] = a.mu_w[q];
        n.mu[1] = a.mu_n[q];
    }
    return n;
}

CAPAZ_HD void adj2_coef_axis(const DevParams& p, const CoefD2& c, bool inner, const double* mu_self, const AdjNb2& m,
                             const AdjNb2& q, double rh, bool grav, double* lP, double* lL, double* lro) {
    for (int ph = 0; ph < 2; ++ph) {
        const double Lc = -c.Xi[ph] * c.ro[ph];
        if (inner) {
            const double gr = grav ? p.g * c.ro[ph] : 0.;
            const double x1 = -((c.P - m.P) * rh) + gr, x2 = -((q.P - c.P) * rh) + gr;
            const DirD d = directed_d(x1, x2, Lc, -m.Xi[ph] * m.ro[ph], -q.Xi[ph] * q.ro[ph], rh);
            lP[ph] += mu_self[ph] * (d.dx2 - d.dx1) * rh;
            lL[ph] += mu_self[ph] * d.dLc;
            if (grav) lro[ph] += mu_self[ph] * (d.dx1 + d.dx2) * p.g;
        }
        if (q.exists && q.interior) {
            const double gr = grav ? p.g * q.ro[ph] : 0.;
            const double x1 = -((q.P - c.P) * rh) + gr;
            const double Lq = -q.Xi[ph] * q.ro[ph];
            const double w1 = upwind_w(x1);
            lP[ph] += q.mu[ph] * ((-(1. - w1) * Lq - w1 * Lc) * rh) * rh;
            lL[ph] += q.mu[ph] * (-(x1 > 0. ? x1 : 0.) * rh);
        }
        if (m.exists && m.interior) {
            const double gr = grav ? p.g * m.ro[ph] : 0.;
            const double x2 = -((c.P - m.P) * rh) + gr;
            const double Lq = -m.Xi[ph] * m.ro[ph];
            const double w2 = upwind_w(x2);
            lP[ph] += m.mu[ph] * (-((w2 * Lq + (1. - w2) * Lc) * rh) * rh);
            lL[ph] += m.mu[ph] * ((x2 < 0. ? x2 : 0.) * rh);
        }
    }
}

template <bool N2 = false>
CAPAZ_HD void adj2_coef_cell(const DevParams& p, const AdjIO& a, int i, int j, int k, int64_t l) {
    const bool inner = interior_newton2(p, i, j, k);
    const CoefD2 c = coef_d2<N2>(p, a, l);
    const double mu_self[2] = {inner ? a.mu_w[l] : 0., inner ? a.mu_n[l] : 0.};
    double lP[2] = {0., 0.}, lL[2] = {0., 0.}, lro[2] = {0., 0.};
    if (!(p.Nx < 2))
        adj2_coef_axis(p, c, inner, mu_self, adj2_nb_at<N2>(p, a, i - 1, j, k), adj2_nb_at<N2>(p, a, i + 1, j, k), p.r_hx,
                       false, lP, lL, lro);
    adj2_coef_axis(p, c, inner, mu_self, adj2_nb_at<N2>(p, a, i, j - 1, k), adj2_nb_at<N2>(p, a, i, j + 1, k), p.r_hy, true,
                   lP, lL, lro);
    if (!(p.Nz < 2))
        adj2_coef_axis(p, c, inner, mu_self, adj2_nb_at<N2>(p, a, i, j, k - 1), adj2_nb_at<N2>(p, a, i, j, k + 1), p.r_hz,
                       false, lP, lL, lro);
    double lPw = 0., lSn = 0., glnK = 0.;
    for (int ph = 0; ph < 2; ++ph) {
        const double lXi = -lL[ph] * c.ro[ph];
        const double lr = lro[ph] - lL[ph] * c.Xi[ph];
        lPw += lP[ph] + lr * c.dro_dP[ph];
        lSn += lXi * c.dXi_dSn[ph];
        glnK += lXi * c.Xi[ph];
    }
    if (inner) {
        lPw += a.dPw[l];
        lSn += a.dSn[l];
    }
    a.lPw[l] = lPw;
    a.lSn[l] = lSn;
    a.g_lnK[l] += glnK;
}

// ---- tangent-linear step -------------------------------------------------------
//
// The forward-mode derivative of the same step: given x_t, its tangent dx_t
// and a parameter direction (d ln K, dm), dx_{t+1}.  Every partial is the one
// the adjoint transposes (CoefD, directed_d with the mean slope at a zero
// driving force, the Newton matrix's columns J - (dM/dx_j) u), so
// <dJ from the tangent> = <gradient, direction> holds to rounding.

struct TanIO {
    const double *Pw, *Sw, *Sn, *Pn, *Pg, *m, *K, *dtm;  // x_t and the parameters
    const double *dPw, *dSw, *dSn, *dPn, *dPg;           // dx_t
    const double *dlnK, *dm;                             // parameter direction
    double *dPw1, *dSw1, *dSn1, *dPn1, *dPg1;            // dx_{t+1}
};

// A cell's coefficients and their tangents.
CAPAZ_HD void coef_tan(const DevParams& p, const TanIO& a, int64_t l, bool inner, Coef& c, Coef& dc) {
    const CoefD d = coef_d(p, a, l, inner);
    const double dP[3] = {a.dPw[l], inner ? a.dPw[l] : a.dPn[l], inner ? a.dPw[l] : a.dPg[l]};
    const double dSw = a.dSw[l], dSn = a.dSn[l], dlnK = a.dlnK[l];
    for (int ph = 0; ph < 3; ++ph) {
        c.P[ph] = d.P[ph];
        c.Xi[ph] = d.Xi[ph];
        c.ro[ph] = d.ro[ph];
        dc.P[ph] = dP[ph];
        dc.Xi[ph] = d.dXi_dSw[ph] * dSw + d.dXi_dSn[ph] * dSn + d.Xi[ph] * dlnK;
        dc.ro[ph] = d.dro_dP[ph] * dP[ph];
    }
}

// Tangent of one cell's flux of phase ph along an axis (c the cell, m / q its
// minus / plus neighbours; grav: the y axis).
CAPAZ_HD double flux_tan(const DevParams& p, int ph, const Coef& c, const Coef& dc, const Coef& m, const Coef& dm,
                         const Coef& q, const Coef& dq, double rh, bool grav) {
    const double gr = grav ? p.g * c.ro[ph] : 0., dgr = grav ? p.g * dc.ro[ph] : 0.;
    const double x1 = -((c.P[ph] - m.P[ph]) * rh) + gr, x2 = -((q.P[ph] - c.P[ph]) * rh) + gr;
    const double dx1 = -((dc.P[ph] - dm.P[ph]) * rh) + dgr, dx2 = -((dq.P[ph] - dc.P[ph]) * rh) + dgr;
    const double Lc = -c.Xi[ph] * c.ro[ph], Lm = -m.Xi[ph] * m.ro[ph], Lq = -q.Xi[ph] * q.ro[ph];
    const double dLc = -(dc.Xi[ph] * c.ro[ph] + c.Xi[ph] * dc.ro[ph]);
    const double dLm = -(dm.Xi[ph] * m.ro[ph] + m.Xi[ph] * dm.ro[ph]);
    const double dLq = -(dq.Xi[ph] * q.ro[ph] + q.Xi[ph] * dq.ro[ph]);
    const DirD d = directed_d(x1, x2, Lc, Lm, Lq, rh);
    return d.dx1 * dx1 + d.dx2 * dx2 + d.dLc * dLc + d.dLm * dLm + d.dLp * dLq;
}

// Tangent of the fixed-iteration Newton: x' = x - M^-1 F, with A and its
// tangent dA; dx (in: of the initial guess, out: of the result).
template <int IT>
CAPAZ_HD void newton3_tangent(const DevParams& p, double P0, double Sw0, double Sn0, const double* A, const double* dA,
                              double* dx) {
    constexpr int NX = IT > 0 ? IT : kMaxAdjIter;
    const int it = IT > 0 ? IT : (p.iterations < kMaxAdjIter ? p.iterations : kMaxAdjIter);
    double P = P0, Sw = Sw0, Sn = Sn0;
    const double rg = p.ro0_g * p.r_P_atm;
#if defined(__HIPCC__)
#pragma unroll
#endif
    for (int k = 0; k < NX; ++k) {
        if (k >= it) break;
        double M[9];
        newton3_matrix(p, P, Sw, Sn, M);
        const double F[3] = {M[1] * Sw - A[0], M[5] * Sn - A[1], -M[7] * (1. - Sw - Sn) - A[2]};
        const double MT[9] = {M[0], M[3], M[6], M[1], M[4], M[7], M[2], M[5], M[8]};
        double u[3];
        solve3_transposed(MT, F, u);
        // (J - (dM/dx_j) u) dx - dA, columns as in newton3_adjoint
        const double u0 = u[0], u1 = u[1], u2 = u[2], gP = rg * P;
        const double cP0 = M[0] - p.c_wb * u1, cP1 = M[3] - p.c_nb * u2, cP2 = M[6] + rg * (u1 + u2);
        const double cW0 = M[1] - p.c_wb * u0, cW2 = -gP + rg * u0;
        const double cN1 = M[5] - (p.c_nb * u0 + p.ro0_n * u1), cN2 = -gP + rg * u0;
        const double r[3] = {cP0 * dx[0] + cW0 * dx[1] - dA[0], cP1 * dx[0] + cN1 * dx[2] - dA[1],
                             cP2 * dx[0] + cW2 * dx[1] + cN2 * dx[2] - dA[2]};
        double du[3];
        solve3_transposed(MT, r, du);
        dx[0] -= du[0];
        dx[1] -= du[1];
        dx[2] -= du[2];
        P -= u0;
        Sw -= u1;
        Sn -= u2;
    }
}

// Interior cell (i, j, k), l its index: dx_{t+1}(l).
template <int IT>
CAPAZ_HD void tan_cell(const DevParams& p, const TanIO& a, int i, int j, int k, int64_t l) {
    Coef c, dc;
    coef_tan(p, a, l, true, c, dc);
    double div[3] = {0., 0., 0.}, ddiv[3] = {0., 0., 0.};
    auto nb = [&](int ii, int jj, int kk, Coef& n, Coef& dn) {
        const int64_t q = (int64_t)ii + (int64_t)jj * p.sy + (int64_t)kk * p.sz;
        coef_tan(p, a, q, interior3(p, ii, jj, kk), n, dn);
    };
    Coef xm = c, xp = c, zm = c, zp = c, ym, yp, dxm{}, dxp{}, dzm{}, dzp{}, dym, dyp;
    if (!(p.Nx < 2)) {
        nb(i - 1, j, k, xm, dxm);
        nb(i + 1, j, k, xp, dxp);
    }
    if (!(p.Nz < 2)) {
        nb(i, j, k - 1, zm, dzm);
        nb(i, j, k + 1, zp, dzp);
    }
    nb(i, j - 1, k, ym, dym);
    nb(i, j + 1, k, yp, dyp);
    for (int ph = 0; ph < 3; ++ph) {
        div[ph] = coef_divergence(p, ph, c, xm, xp, ym, yp, zm, zp);
        if (!(p.Nx < 2)) ddiv[ph] += flux_tan(p, ph, c, dc, xm, dxm, xp, dxp, p.r_hx, false);
        ddiv[ph] += flux_tan(p, ph, c, dc, ym, dym, yp, dyp, p.r_hy, true);
        if (!(p.Nz < 2)) ddiv[ph] += flux_tan(p, ph, c, dc, zm, dzm, zp, dzp, p.r_hz, false);
    }
    const double Sw = a.Sw[l], Sn = a.Sn[l], Sg = 1. - Sw - Sn, dtm = a.dtm[l];
    const double dSw = a.dSw[l], dSn = a.dSn[l];
    const double ddtm = -dtm / a.m[l] * a.dm[l];
    const double A[3] = {c.ro[0] * Sw - dtm * div[0], c.ro[1] * Sn - dtm * div[1], c.ro[2] * Sg - dtm * div[2]};
    const double dA[3] = {dc.ro[0] * Sw + c.ro[0] * dSw - ddtm * div[0] - dtm * ddiv[0],
                          dc.ro[1] * Sn + c.ro[1] * dSn - ddtm * div[1] - dtm * ddiv[1],
                          dc.ro[2] * Sg - c.ro[2] * (dSw + dSn) - ddtm * div[2] - dtm * ddiv[2]};
    double dx[3] = {a.dPw[l], dSw, dSn};
    newton3_tangent<IT>(p, a.Pw[l], Sw, Sn, A, dA, dx);
    a.dPw1[l] = dx[0];
    a.dSw1[l] = dx[1];
    a.dSn1[l] = dx[2];
    a.dPn1[l] = 0.;  // interior P_n, P_g follow P_w and are never read
    a.dPg1[l] = 0.;
}

// Boundary cell (i, j, k) from its inward cell's new tangent (Border_S /
// border_p3, three-phase.cpp:345-427).
CAPAZ_HD void tan_boundary(const DevParams& p, const TanIO& a, int i, int j, int k) {
    const int64_t l = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz, l1 = inward(p, i, j, k);
    double dw = 1., dn = 1., dg = 1.;
    if (j == 0) {
        dw = p.r_bp_w_top_den;
        dn = p.r_bp_n_top_den;
        dg = p.r_bp_g_top_den;
    } else if (j == p.Ny - 1) {
        dw = p.r_bp_w_bot_den;
        dn = p.r_bp_n_bot_den;
        dg = p.r_bp_g_bot_den;
    }
    const double d1 = a.dPw1[l1];
    a.dPw1[l] = dw * d1;
    a.dPn1[l] = dn * d1;
    a.dPg1[l] = dg * d1;
    const bool injected = j == 0 && p.source > 0;
    a.dSw1[l] = injected ? 0. : a.dSw1[l1];
    a.dSn1[l] = injected ? 0. : a.dSn1[l1];
}

}  // namespace capaz
