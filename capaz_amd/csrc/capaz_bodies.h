// Per-cell bodies of the entry-point kernels and of the fused step kernel.
// Each body handles one cell `l` and is called by a HIP kernel (one thread per
// cell) or, in the host build, by a plain loop.  Field arrays are SoA, x
// fastest.
#pragma once

#include "capaz_cell.h"

namespace capaz {

enum FieldId {
    F_S_n = 0, F_S_w, F_P_w, F_P_n, F_ro_w, F_ro_n, F_ux_w, F_uy_w, F_uz_w, F_ux_n, F_uy_n, F_uz_n,
    F_Xi_w, F_Xi_n, F_roS_w, F_roS_w_old, F_roS_n, F_roS_n_old, F_m, F_K,
    F_S_g, F_P_g, F_ro_g, F_ux_g, F_uy_g, F_uz_g, F_Xi_g, F_roS_g, F_roS_g_old,
    F_T, F_H_w, F_H_n, F_H_g, F_H_r, F_E, F_E_new, F_COUNT
};

struct DevFields {
    double* f[F_COUNT];
};

CAPAZ_HD void decompose(const DevParams& p, int64_t l, int& i, int& j, int& k) {
    int64_t jk = l / p.Nx;
    i = (int)(l - jk * p.Nx);
    k = (int)(jk / p.Ny);
    j = (int)(jk - (int64_t)k * p.Ny);
}

// ---- S_calculation: assign_S (cpu.cpp:171-182), all cells ----------------------
CAPAZ_HD void body_S(const DevParams& p, const DevFields& F, int64_t l) {
    if (p.model == 3)
        F.f[F_S_g][l] = 1. - F.f[F_S_w][l] - F.f[F_S_n][l];
    else
        F.f[F_S_w][l] = 1. - F.f[F_S_n][l];
}

// ---- ro_P_Xi_calculation: assign_P_Xi + assign_ro (cpu.cpp:3-15), all cells -----
CAPAZ_HD void body_P_Xi_ro(const DevParams& p, const DevFields& F, int64_t l) {
    int i, j, k;
    decompose(p, l, i, j, k);
    double K = F.f[F_K][l], Pw = F.f[F_P_w][l];
    if (p.model == 2) {
        double kw, kn, Pk;
        relperm2(p, F.f[F_S_n][l], kw, kn, Pk);
        double Pn = Pw + Pk;
        F.f[F_P_n][l] = Pn;
        if (p.energy) {  // the configurations' energy closure: mu(T), energy.cpp's ro(P, T)
            const double T = F.f[F_T][l];
            F.f[F_Xi_w][l] = xi_var(K, kw, mu_cfg(p, p.mu_w, T));
            F.f[F_Xi_n][l] = xi_var(K, kn, mu_cfg(p, p.mu_n, T));
            F.f[F_ro_w][l] = ro_w_T(p, Pw, T);
            F.f[F_ro_n][l] = ro_n_T(p, Pn, T);
            return;
        }
        F.f[F_Xi_w][l] = xi_const(K, kw, p.mu_w, p.r_mu_w);
        F.f[F_Xi_n][l] = xi_const(K, kn, p.mu_n, p.r_mu_n);
        F.f[F_ro_w][l] = ro_w_of(p, Pw);
        F.f[F_ro_n][l] = ro_n_of(p, Pn);
        return;
    }
    double kw, kn, kg;
    relperm3(p, F.f[F_S_w][l], F.f[F_S_n][l], kw, kn, kg);
    double Pn, Pg;
    if (interior3(p, i, j, k)) {
        Pn = Pw;  // P_w + Pk_nw with Pk_nw == 0 (three-phase.cpp:230-231)
        Pg = Pw;
        F.f[F_P_n][l] = Pn;
        F.f[F_P_g][l] = Pg;
    } else {
        Pn = F.f[F_P_n][l];
        Pg = F.f[F_P_g][l];
    }
    if (p.energy) {
        double T = F.f[F_T][l], mw, mn, mg;
        if (p.ecfg) {
            mw = mu_cfg(p, p.mu_w, T);
            mn = mu_cfg(p, p.mu_n, T);
            mg = mu_cfg(p, p.mu_g, T);
        } else {
            mu_of_T(T, mw, mn, mg);
        }
        F.f[F_Xi_w][l] = xi_var(K, kw, mw);
        F.f[F_Xi_n][l] = xi_var(K, kn, mn);
        F.f[F_Xi_g][l] = xi_var(K, kg, mg);
        F.f[F_ro_w][l] = ro_w_T(p, Pw, T);
        F.f[F_ro_n][l] = ro_n_T(p, Pn, T);
        F.f[F_ro_g][l] = ro_g_T(p, Pg, T);
    } else {
        F.f[F_Xi_w][l] = xi_const(K, kw, p.mu_w, p.r_mu_w);
        F.f[F_Xi_n][l] = xi_const(K, kn, p.mu_n, p.r_mu_n);
        F.f[F_Xi_g][l] = xi_const(K, kg, p.mu_g, p.r_mu_g);
        F.f[F_ro_w][l] = ro_w_of(p, Pw);
        F.f[F_ro_n][l] = ro_n_of(p, Pn);
        F.f[F_ro_g][l] = ro_g_of(p, Pg);
    }
}

// ---- u_calculation: assign_u (cpu.cpp:346-477), all cells ------------------------
CAPAZ_HD double u_axis(const double* Xi, const double* P, const double* ro, int64_t l, int64_t s, int idx, int n,
                       double h, double rh, double h2, double rh2, bool grav, double g) {
    if (n <= 2) return 0.;
    double d;
    if (idx == 0)
        d = exact_div(P[l + s] - P[l], h, rh);
    else if (idx == n - 1)
        d = exact_div(P[l] - P[l - s], h, rh);
    else
        d = exact_div(P[l + s] - P[l - s], h2, rh2);
    return grav ? Xi[l] * (d - ro[l] * g) : Xi[l] * d;
}

CAPAZ_HD void body_u(const DevParams& p, const DevFields& F, int64_t l) {
    int i, j, k;
    decompose(p, l, i, j, k);
    const int ph = p.model == 3 ? 3 : 2;
    const int xi[3] = {F_Xi_w, F_Xi_n, F_Xi_g}, pr[3] = {F_P_w, F_P_n, F_P_g}, ro[3] = {F_ro_w, F_ro_n, F_ro_g};
    const int ux[3] = {F_ux_w, F_ux_n, F_ux_g}, uy[3] = {F_uy_w, F_uy_n, F_uy_g}, uz[3] = {F_uz_w, F_uz_n, F_uz_g};
    for (int a = 0; a < ph; ++a) {
        const double *X = F.f[xi[a]], *P = F.f[pr[a]], *R = F.f[ro[a]];
        F.f[ux[a]][l] = u_axis(X, P, R, l, 1, i, p.Nx, p.hx, p.r_hx, p.two_hx, p.r_two_hx, false, p.g);
        F.f[uy[a]][l] = u_axis(X, P, R, l, p.sy, j, p.Ny, p.hy, p.r_hy, p.two_hy, p.r_two_hy, true, p.g);
        F.f[uz[a]][l] = u_axis(X, P, R, l, p.sz, k, p.Nz, p.hz, p.r_hz, p.two_hz, p.r_two_hz, false, p.g);
    }
}

// ---- roS_calculation: assign_roS_nr (cpu.cpp:619-731), interior cells ----------------

// Flux of one phase along one axis from the field arrays (cpu.cpp:661-705).
CAPAZ_HD double field_flux(const DevParams& p, const double* P, const double* Xi, const double* ro, int64_t l,
                           int64_t s, int axis) {
    double h = axis == 0 ? p.hx : (axis == 1 ? p.hy : p.hz);
    double rh = axis == 0 ? p.r_hx : (axis == 1 ? p.r_hy : p.r_hz);
    double dp = exact_div(P[l + s] - P[l], h, rh);
    double dm = exact_div(P[l] - P[l - s], h, rh);
    if (axis == 1) return flux_y(p, dm, dp, Xi[l], ro[l], Xi[l - s], ro[l - s], Xi[l + s], ro[l + s]);
    return flux_x(dm, dp, Xi[l], ro[l], Xi[l - s], ro[l - s], Xi[l + s], ro[l + s], h, rh);
}

CAPAZ_HD double phase_divergence(const DevParams& p, const DevFields& F, int a, int64_t l) {
    const int xi[3] = {F_Xi_w, F_Xi_n, F_Xi_g}, pr[3] = {F_P_w, F_P_n, F_P_g}, ro[3] = {F_ro_w, F_ro_n, F_ro_g};
    const double *X = F.f[xi[a]], *P = F.f[pr[a]], *R = F.f[ro[a]];
    double fz = p.Nz < 2 ? 0. : field_flux(p, P, X, R, l, p.sz, 2);
    double fx = p.Nx < 2 ? 0. : field_flux(p, P, X, R, l, 1, 0);
    double fy = field_flux(p, P, X, R, l, p.sy, 1);
    // (-q + fx + fy + fz) for water/NAPL, (q + ...) for gas, with q == 0
    // (cpu.cpp:707-708, 723): -0. + fx == fx exactly; 0. + fx kept for gas.
    double s = (a == 2 ? 0. + fx : fx) + fy;
    return s + fz;
}

CAPAZ_HD void body_roS(const DevParams& p, const DevFields& F, int64_t l) {
    int i, j, k;
    decompose(p, l, i, j, k);
    if (!interior3(p, i, j, k)) return;
    double m = F.f[F_m][l];
    if (!m) return;
    double Sn = F.f[F_S_n][l];
    double roSw, roSn = F.f[F_ro_n][l] * Sn, roSg = 0.;
    if (p.model == 3) {
        double Sw = F.f[F_S_w][l];
        roSw = F.f[F_ro_w][l] * Sw;
        roSg = F.f[F_ro_g][l] * (1. - Sw - Sn);
    } else {
        roSw = F.f[F_ro_w][l] * (1. - Sn);
    }
    double dtm = p.dt / m;
    double A1 = roSw - dtm * phase_divergence(p, F, 0, l);
    double A2 = roSn - dtm * phase_divergence(p, F, 1, l);
    F.f[F_roS_w_old][l] = roSw;
    F.f[F_roS_n_old][l] = roSn;
    F.f[F_roS_w][l] = A1;
    F.f[F_roS_n][l] = A2;
    if (p.model == 3) {
        double A3 = roSg - dtm * phase_divergence(p, F, 2, l);
        F.f[F_roS_g_old][l] = roSg;
        F.f[F_roS_g][l] = A3;
    }
}

// ---- energy: H_E_current_calculation (cpu.cpp:38-50; energy.cpp:132-143,317-325) ----
CAPAZ_HD void body_HE(const DevParams& p, const DevFields& F, int64_t l) {
    double T = F.f[F_T][l];
    double Hw = H_w_of(p, T), Hn = H_n_T(p, T), Hr = H_r_of(p, T);
    F.f[F_H_w][l] = Hw;
    F.f[F_H_n][l] = Hn;
    F.f[F_H_r][l] = Hr;
    double m = F.f[F_m][l];
    if (p.model == 2) {  // assign_E_current without the gas term
        F.f[F_E][l] = (m * (F.f[F_S_w][l] * (F.f[F_ro_w][l] * Hw - F.f[F_P_w][l]) +
                            F.f[F_S_n][l] * (F.f[F_ro_n][l] * Hn - F.f[F_P_n][l])) +
                       (1. - m) * (p.ro_r * Hr - F.f[F_P_w][l]));
        return;
    }
    double Hg = H_g_T(p, T);
    F.f[F_H_g][l] = Hg;
    F.f[F_E][l] = (m * (F.f[F_S_w][l] * (F.f[F_ro_w][l] * Hw - F.f[F_P_w][l]) +
                        F.f[F_S_n][l] * (F.f[F_ro_n][l] * Hn - F.f[F_P_n][l]) +
                        F.f[F_S_g][l] * (F.f[F_ro_g][l] * Hg - F.f[F_P_g][l])) +
                   (1. - m) * (p.ro_r * Hr - F.f[F_P_w][l]));
}

// ---- energy: E_calculation -> assign_E_new (energy.cpp:235-340), interior cells ---
CAPAZ_HD double lambda_eff_at(const DevParams& p, const DevFields& F, int64_t l) {
    if (p.ecfg) return p.lam_eff;
    return lambda_eff(F.f[F_m][l], F.f[F_S_w][l], F.f[F_S_n][l], F.f[F_S_g][l], F.f[F_T][l]);
}

// lambda_eff of a cell, computed from the fields (the energy step's tiled E
// pass reads the same values from an array instead).
struct LamOfFields {
    const DevParams& p;
    const DevFields& F;
    CAPAZ_HD double operator()(int64_t l) const { return lambda_eff_at(p, F, l); }
};

template <class Lam>
CAPAZ_HD double conduction(const DevFields& F, const Lam& lam, int64_t l, int64_t s, double h2, double rh2) {
    const double* T = F.f[F_T];
    return exact_div(lam(l + s) * T[l + s] - 2 * lam(l) * T[l] + lam(l - s) * T[l - s], h2, rh2);
}

CAPAZ_HD double advection(const DevParams& p, const DevFields& F, const double* uw, const double* un,
                          const double* ug, int64_t a, int64_t b) {
    const double *rw = F.f[F_ro_w], *rn = F.f[F_ro_n];
    const double *Hw = F.f[F_H_w], *Hn = F.f[F_H_n];
    const double s = rw[a] * Hw[a] * uw[a] - rw[b] * Hw[b] * uw[b] + rn[a] * Hn[a] * un[a] - rn[b] * Hn[b] * un[b];
    if (p.model == 2) return s;  // no gas terms
    const double *rg = F.f[F_ro_g], *Hg = F.f[F_H_g];
    return s + rg[a] * Hg[a] * ug[a] - rg[b] * Hg[b] * ug[b];
}

template <class Lam>
CAPAZ_HD void body_E_with(const DevParams& p, const DevFields& F, const Lam& lam, int64_t l) {
    int i, j, k;
    decompose(p, l, i, j, k);
    if (!interior3(p, i, j, k)) return;
    double T_flow = 0;
    if (p.Nx > 2) T_flow += conduction(F, lam, l, 1, p.hx2, p.r_hx2);
    if (p.Ny > 2) T_flow += conduction(F, lam, l, p.sy, p.hy2, p.r_hy2);
    if (p.Nz > 2) {
        const double tz = conduction(F, lam, l, p.sz, p.hz2, p.r_hz2);
        if (p.ecfg)
            T_flow += tz;
        else
            T_flow = tz;  // '=' as energy.cpp:257
    }
    double E_flow = 0;
    if (p.Nx > 2)
        E_flow += exact_div(advection(p, F, F.f[F_ux_w], F.f[F_ux_n], F.f[F_ux_g], l + 1, l - 1), p.two_hx,
                            p.r_two_hx);
    if (p.Ny > 2)
        E_flow += exact_div(advection(p, F, F.f[F_uy_w], F.f[F_uy_n], F.f[F_uy_g], l + p.sy, l - p.sy), p.two_hy,
                            p.r_two_hy);
    if (p.Nz > 2) {  // the reference advects with uy here (energy.cpp:300-305)
        const int w = p.ecfg ? F_uz_w : F_uy_w, n = p.ecfg ? F_uz_n : F_uy_n, g = p.ecfg ? F_uz_g : F_uy_g;
        E_flow += exact_div(advection(p, F, F.f[w], F.f[n], F.f[g], l + p.sz, l - p.sz), p.two_hz, p.r_two_hz);
    }
    double Q_hw = 0, Q_hr = 0;
    F.f[F_E_new][l] = F.f[F_E][l] + p.dt * (T_flow + Q_hw + Q_hr - E_flow);
}

CAPAZ_HD void body_E(const DevParams& p, const DevFields& F, int64_t l) { body_E_with(p, F, LamOfFields{p, F}, l); }

// ---- P_S_calculation -> Newton (cpu.cpp:80-89), interior cells -------------------------
// The energy model's Newton update of interior cell l without the pivot
// search: writes the new state and returns true, or writes nothing and
// returns false when a pivot needs the search (body_newton then does the cell).
CAPAZ_HD bool body_newton_energy_nopivot(const DevParams& p, const DevFields& F, int64_t l) {
    double P = F.f[F_P_w][l], Sw = F.f[F_S_w][l], Sn = F.f[F_S_n][l];
    double Sg = F.f[F_S_g][l], T = F.f[F_T][l];
    double Hw = F.f[F_H_w][l], Hn = F.f[F_H_n][l], Hg = F.f[F_H_g][l], Hr = F.f[F_H_r][l];
    const double E_new = F.f[F_E_new][l];
    if (!newton_energy_t<true>(p, Sw, Sn, Sg, P, T, Hw, Hn, Hg, Hr, F.f[F_m][l], F.f[F_roS_w][l], F.f[F_roS_n][l],
                               F.f[F_roS_g][l], E_new))
        return false;
    F.f[F_S_g][l] = Sg;
    F.f[F_T][l] = T;
    F.f[F_H_w][l] = Hw;
    F.f[F_H_n][l] = Hn;
    F.f[F_H_g][l] = Hg;
    F.f[F_H_r][l] = Hr;
    F.f[F_E][l] = E_new;
    F.f[F_P_w][l] = P;
    F.f[F_S_w][l] = Sw;
    F.f[F_S_n][l] = Sn;
    return true;
}

// The two-phase energy model's Newton update of interior cell l without the
// pivot search (false, nothing written: the cell needs the search).
CAPAZ_HD bool body_newton_energy2_nopivot(const DevParams& p, const DevFields& F, int64_t l) {
    double P = F.f[F_P_w][l], Sw = F.f[F_S_w][l], Sn = F.f[F_S_n][l], T = F.f[F_T][l];
    double Hw = F.f[F_H_w][l], Hn = F.f[F_H_n][l], Hr = F.f[F_H_r][l];
    const double E_new = F.f[F_E_new][l];
    if (!newton_energy2_t<true>(p, Sw, Sn, P, T, Hw, Hn, Hr, F.f[F_m][l], F.f[F_roS_w][l], F.f[F_roS_n][l], E_new))
        return false;
    F.f[F_T][l] = T;
    F.f[F_H_w][l] = Hw;
    F.f[F_H_n][l] = Hn;
    F.f[F_H_r][l] = Hr;
    F.f[F_E][l] = E_new;
    F.f[F_P_w][l] = P;
    F.f[F_S_w][l] = Sw;
    F.f[F_S_n][l] = Sn;
    return true;
}

CAPAZ_HD void body_newton(const DevParams& p, const DevFields& F, int64_t l) {
    int i, j, k;
    decompose(p, l, i, j, k);
    if (p.model == 2 && p.energy) {
        if (!interior_newton2(p, i, j, k)) return;
        double P = F.f[F_P_w][l], Sw = F.f[F_S_w][l], Sn = F.f[F_S_n][l], T = F.f[F_T][l];
        double Hw = F.f[F_H_w][l], Hn = F.f[F_H_n][l], Hr = F.f[F_H_r][l];
        const double E_new = F.f[F_E_new][l];
        newton_energy2_t<false>(p, Sw, Sn, P, T, Hw, Hn, Hr, F.f[F_m][l], F.f[F_roS_w][l], F.f[F_roS_n][l], E_new);
        F.f[F_T][l] = T;
        F.f[F_H_w][l] = Hw;
        F.f[F_H_n][l] = Hn;
        F.f[F_H_r][l] = Hr;
        F.f[F_E][l] = E_new;
        F.f[F_P_w][l] = P;
        F.f[F_S_w][l] = Sw;
        F.f[F_S_n][l] = Sn;
        return;
    }
    if (p.model == 2) {
        if (!interior_newton2(p, i, j, k)) return;
        double P = F.f[F_P_w][l], Sn = F.f[F_S_n][l];
        newton2(p, P, Sn, F.f[F_roS_w][l], F.f[F_roS_n][l]);
        F.f[F_P_w][l] = P;
        F.f[F_S_n][l] = Sn;
        return;
    }
    if (!interior3(p, i, j, k)) return;
    double P = F.f[F_P_w][l], Sw = F.f[F_S_w][l], Sn = F.f[F_S_n][l];
    if (p.energy) {
        double Sg = F.f[F_S_g][l], T = F.f[F_T][l];
        double Hw = F.f[F_H_w][l], Hn = F.f[F_H_n][l], Hg = F.f[F_H_g][l], Hr = F.f[F_H_r][l];
        double E_new = F.f[F_E_new][l];
        newton_energy(p, Sw, Sn, Sg, P, T, Hw, Hn, Hg, Hr, F.f[F_m][l], F.f[F_roS_w][l], F.f[F_roS_n][l],
                      F.f[F_roS_g][l], E_new);
        F.f[F_S_g][l] = Sg;
        F.f[F_T][l] = T;
        F.f[F_H_w][l] = Hw;
        F.f[F_H_n][l] = Hn;
        F.f[F_H_g][l] = Hg;
        F.f[F_H_r][l] = Hr;
        F.f[F_E][l] = E_new;
    } else {
        newton3(p, P, Sw, Sn, F.f[F_roS_w][l], F.f[F_roS_n][l], F.f[F_roS_g][l]);
    }
    F.f[F_P_w][l] = P;
    F.f[F_S_w][l] = Sw;
    F.f[F_S_n][l] = Sn;
}



// ---- boundary_conditions (cpu.cpp:91-104), boundary cells -------------------------------
CAPAZ_HD void body_boundary(const DevParams& p, const DevFields& F, int64_t l) {
    int i, j, k;
    decompose(p, l, i, j, k);
    if (!is_boundary(p, i, j, k)) return;
    int64_t l1 = inward(p, i, j, k);
    if (p.model == 2) {  // two-phase.cpp:163-215
        double Sn1 = F.f[F_S_n][l1];
        double Sn = Sn1;
        if (j == 0 && p.source > 0 && in_source_window(p, i, k)) Sn = p.S_n_gr;
        F.f[F_S_n][l] = Sn;
        double P1 = F.f[F_P_w][l1], Pw;
        if ((j != 0) && (j != p.Ny - 1))
            Pw = P1;
        else if (j == 0)
            Pw = p.P_atm;
        else if (p.P_bottom > 0)  // the configurations' fixed bottom pressure
            Pw = p.P_bottom;
        else  // ro_eff_gdy (main.cpp:255-267)
            Pw = P1 + (F.f[F_ro_n][l1] * Sn1 + F.f[F_ro_w][l1] * (1 - Sn1)) * F.f[F_m][l1] * p.g * p.hy;
        F.f[F_P_w][l] = Pw;
        if (p.energy) F.f[F_T][l] = border_T_cfg(p, j, F.f[F_T][l1]);
        return;
    }
    // three-phase.cpp:345-427
    if (j == 0 && p.source > 0) {
        F.f[F_S_w][l] = p.S_w_gr;
        F.f[F_S_n][l] = p.S_n_gr;
    } else {
        F.f[F_S_w][l] = F.f[F_S_w][l1];
        F.f[F_S_n][l] = F.f[F_S_n][l1];
    }
    double Pw, Pn, Pg;
    border_p3(p, j, F.f[F_P_w][l1], Pw, Pn, Pg);
    F.f[F_P_w][l] = Pw;
    F.f[F_P_n][l] = Pn;
    F.f[F_P_g][l] = Pg;
    if (p.energy && p.ecfg) {
        F.f[F_T][l] = border_T_cfg(p, j, F.f[F_T][l1]);
    } else if (p.energy) {  // Border_T (energy.cpp:343-366)
        double T;
        if (j == 0)
            T = 400;
        else if (j == p.Ny - 1)
            T = 273;
        else
            T = F.f[F_T][l1];
        F.f[F_T][l] = T;
    }
}

// ---- fused energy step (three-phase + energy) ----------------------------------------------
//
// The entry-point sequence S, P_Xi_ro, u, roS, HE, E, Newton, boundary in six
// passes, each cell's arithmetic that of the entry-point bodies above:
//   A (all cells)      S, P_Xi_ro and HE of the cell (S_g, interior P_n/P_g,
//                      Xi, ro, H, E in place), and the cell's lambda_eff into
//                      a work array: E's conduction reads it at 7 points per
//                      cell instead of evaluating it (a pow) 9 times;
//   U (boundary cells) u;
//   C (interior)       roS and u on the fused kernel's z tiles, from the
//                      coefficients A stored (k_roS_z, capaz_fused.h);
//   E (interior)       E_new on z tiles, from lambda_eff * T and the advected
//                      products staged in LDS (k_energy_e_z, capaz_fused.h);
//   N (interior)       the Newton update;
//   D (boundary cells) the boundary.
// E stays apart from C: fused per cell, their ~30 arrays read at 7 points
// ran slower than apart; the Newton update holds ~200 registers, which would
// cut the stencils' occupancy.

CAPAZ_HD void body_energy_a(const DevParams& p, const DevFields& F, double* lam, int64_t l) {
    body_S(p, F, l);
    body_P_Xi_ro(p, F, l);
    body_HE(p, F, l);
    lam[l] = lambda_eff_at(p, F, l);
}

// ---- fused step (two- and three-phase without energy) -------------------------------------
//
// One thread advances one interior cell c through assign_P_Xi/assign_ro of c
// and its six neighbours, assign_roS_nr, Newton, and then writes every
// boundary cell b whose inward neighbour (cpu.cpp:107-141) is c, from c's new
// state (Border_S/Border_P).  Reads the current primary state, writes the
// next one; intermediate fields are not stored.

struct Coef {
    double P[3], Xi[3], ro[3];
};

struct StepIO {
    const double *Pw, *Sw, *Sn, *Pn, *Pg, *m, *K;  // current state (Sw unused for two phases)
    const double* dtm;                             // dt / m per cell (fused z kernel)
    double *Pw1, *Sw1, *Sn1, *Pn1, *Pg1;           // next state
    // Xi and ro per phase as stored by P_Xi_ro (read by the fused z kernel's
    // stored-coefficient variant only, the energy step's roS pass)
    const double *Xi[3] = {}, *ro[3] = {};
    // non-NULL: the step also writes the z exchange planes of its new state
    // here (fused_pack)
    double* pack = nullptr;
};

// assign_P_Xi + assign_ro of one cell from the current primary state.
CAPAZ_HD Coef cell_coef(const DevParams& p, const StepIO& io, int64_t l, bool inner) {
    Coef c;
    double Pw = io.Pw[l], K = io.K[l];
    if (p.model == 2) {
        double kw, kn, Pk;
        relperm2(p, io.Sn[l], kw, kn, Pk);
        double Pn = Pw + Pk;
        c.P[0] = Pw;
        c.P[1] = Pn;
        c.Xi[0] = xi_const(K, kw, p.mu_w, p.r_mu_w);
        c.Xi[1] = xi_const(K, kn, p.mu_n, p.r_mu_n);
        c.ro[0] = ro_w_of(p, Pw);
        c.ro[1] = ro_n_of(p, Pn);
        c.P[2] = c.Xi[2] = c.ro[2] = 0.;
        return c;
    }
    double kw, kn, kg;
    relperm3(p, io.Sw[l], io.Sn[l], kw, kn, kg);
    double Pn = inner ? Pw : io.Pn[l];
    double Pg = inner ? Pw : io.Pg[l];
    c.P[0] = Pw;
    c.P[1] = Pn;
    c.P[2] = Pg;
    c.Xi[0] = xi_const(K, kw, p.mu_w, p.r_mu_w);
    c.Xi[1] = xi_const(K, kn, p.mu_n, p.r_mu_n);
    c.Xi[2] = xi_const(K, kg, p.mu_g, p.r_mu_g);
    c.ro[0] = ro_w_of(p, Pw);
    c.ro[1] = ro_n_of(p, Pn);
    c.ro[2] = ro_g_of(p, Pg);
    return c;
}

CAPAZ_HD bool cell_inner(const DevParams& p, int i, int j, int k) {
    return p.model == 2 ? interior_newton2(p, i, j, k) : interior3(p, i, j, k);
}

// The divergence of one phase a from the coefficients of c and its
// neighbours (order as cpu.cpp:651-708: z, x, y).
CAPAZ_HD double coef_divergence(const DevParams& p, int a, const Coef& c, const Coef& xm, const Coef& xp,
                                const Coef& ym, const Coef& yp, const Coef& zm, const Coef& zp) {
    double fz = 0., fx = 0.;
    if (!(p.Nz < 2)) {
        double dp = exact_div(zp.P[a] - c.P[a], p.hz, p.r_hz), dm = exact_div(c.P[a] - zm.P[a], p.hz, p.r_hz);
        fz = flux_x(dm, dp, c.Xi[a], c.ro[a], zm.Xi[a], zm.ro[a], zp.Xi[a], zp.ro[a], p.hz, p.r_hz);
    }
    if (!(p.Nx < 2)) {
        double dp = exact_div(xp.P[a] - c.P[a], p.hx, p.r_hx), dm = exact_div(c.P[a] - xm.P[a], p.hx, p.r_hx);
        fx = flux_x(dm, dp, c.Xi[a], c.ro[a], xm.Xi[a], xm.ro[a], xp.Xi[a], xp.ro[a], p.hx, p.r_hx);
    }
    double dp = exact_div(yp.P[a] - c.P[a], p.hy, p.r_hy), dm = exact_div(c.P[a] - ym.P[a], p.hy, p.r_hy);
    double fy = flux_y(p, dm, dp, c.Xi[a], c.ro[a], ym.Xi[a], ym.ro[a], yp.Xi[a], yp.ro[a]);
    double s = (a == 2 ? 0. + fx : fx) + fy;
    return s + fz;
}

// Boundary cells whose inward neighbour is the interior cell (i, j, k): the
// offsets -1/+1 along an axis exist when the cell is next to that face.
CAPAZ_HD void boundary_epilogue(const DevParams& p, const StepIO& io, int i, int j, int k, double Pw, double Sw,
                                double Sn, double row_c, double ron_c, double m_c) {
    int ox[3] = {0, 0, 0}, oy[3] = {0, 0, 0}, oz[3] = {0, 0, 0};
    int nx = 1, ny = 1, nz = 1;
    if (p.Nx > 2 && i == 1) ox[nx++] = -1;
    if (p.Nx > 2 && i == p.Nx - 2) ox[nx++] = +1;
    if (j == 1) oy[ny++] = -1;
    if (j == p.Ny - 2) oy[ny++] = +1;
    if (p.Nz > 2 && k == 1) oz[nz++] = -1;
    if (p.Nz > 2 && k == p.Nz - 2) oz[nz++] = +1;
    if (nx == 1 && ny == 1 && nz == 1) return;
    double ro_eff = 0.;
    if (p.model == 2) ro_eff = (ron_c * Sn + row_c * (1 - Sn)) * m_c * p.g * p.hy;
    for (int a = 0; a < nz; ++a)
        for (int b = 0; b < ny; ++b)
            for (int c = 0; c < nx; ++c) {
                if (a == 0 && b == 0 && c == 0) continue;
                int bi = i + ox[c], bj = j + oy[b], bk = k + oz[a];
                int64_t lb = (int64_t)bi + (int64_t)bj * p.sy + (int64_t)bk * p.sz;
                if (p.model == 2) {
                    double Snb = Sn;
                    if (bj == 0 && p.source > 0 && in_source_window(p, bi, bk)) Snb = p.S_n_gr;
                    io.Sn1[lb] = Snb;
                    double Pb;
                    if ((bj != 0) && (bj != p.Ny - 1))
                        Pb = Pw;
                    else if (bj == 0)
                        Pb = p.P_atm;
                    else if (p.P_bottom > 0)
                        Pb = p.P_bottom;
                    else
                        Pb = Pw + ro_eff;
                    io.Pw1[lb] = Pb;
                } else {
                    if (bj == 0 && p.source > 0) {
                        io.Sw1[lb] = p.S_w_gr;
                        io.Sn1[lb] = p.S_n_gr;
                    } else {
                        io.Sw1[lb] = Sw;
                        io.Sn1[lb] = Sn;
                    }
                    double Pwb, Pnb, Pgb;
                    border_p3(p, bj, Pw, Pwb, Pnb, Pgb);
                    io.Pw1[lb] = Pwb;
                    io.Pn1[lb] = Pnb;
                    io.Pg1[lb] = Pgb;
                }
            }
}

// One fused time step of interior cell l (no-op for boundary cells, which
// the epilogue of their inward neighbour writes).
CAPAZ_HD void body_fused_step(const DevParams& p, const StepIO& io, int64_t l) {
    int i, j, k;
    decompose(p, l, i, j, k);
    if (!cell_inner(p, i, j, k)) return;
    const int64_t sx = 1;
    Coef c = cell_coef(p, io, l, true);
    Coef xm = c, xp = c, ym, yp, zm = c, zp = c;
    if (!(p.Nx < 2)) {
        xm = cell_coef(p, io, l - sx, cell_inner(p, i - 1, j, k));
        xp = cell_coef(p, io, l + sx, cell_inner(p, i + 1, j, k));
    }
    ym = cell_coef(p, io, l - p.sy, cell_inner(p, i, j - 1, k));
    yp = cell_coef(p, io, l + p.sy, cell_inner(p, i, j + 1, k));
    if (!(p.Nz < 2)) {
        zm = cell_coef(p, io, l - p.sz, cell_inner(p, i, j, k - 1));
        zp = cell_coef(p, io, l + p.sz, cell_inner(p, i, j, k + 1));
    }
    double m = io.m[l];
    double Pw = io.Pw[l], Sn = io.Sn[l];
    double dtm = p.dt / m;
    if (p.model == 2) {
        double roSw = c.ro[0] * (1. - Sn), roSn = c.ro[1] * Sn;
        double A1 = roSw - dtm * coef_divergence(p, 0, c, xm, xp, ym, yp, zm, zp);
        double A2 = roSn - dtm * coef_divergence(p, 1, c, xm, xp, ym, yp, zm, zp);
        newton2(p, Pw, Sn, A1, A2);
        io.Pw1[l] = Pw;
        io.Sn1[l] = Sn;
        boundary_epilogue(p, io, i, j, k, Pw, 0., Sn, c.ro[0], c.ro[1], m);
        return;
    }
    double Sw = io.Sw[l];
    double roSw = c.ro[0] * Sw, roSg = c.ro[2] * (1. - Sw - Sn), roSn = c.ro[1] * Sn;
    double A1 = roSw - dtm * coef_divergence(p, 0, c, xm, xp, ym, yp, zm, zp);
    double A2 = roSn - dtm * coef_divergence(p, 1, c, xm, xp, ym, yp, zm, zp);
    double A3 = roSg - dtm * coef_divergence(p, 2, c, xm, xp, ym, yp, zm, zp);
    newton3(p, Pw, Sw, Sn, A1, A2, A3);
    io.Pw1[l] = Pw;
    io.Sw1[l] = Sw;
    io.Sn1[l] = Sn;
    boundary_epilogue(p, io, i, j, k, Pw, Sw, Sn, c.ro[0], c.ro[1], m);
}

}  // namespace capaz
