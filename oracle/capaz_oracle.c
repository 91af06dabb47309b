/* The oracle: CPU restatement of capaz's explicit filtration step.
 *
 * TEST INFRASTRUCTURE ONLY -- the reference that the HIP backend is compared
 * with.  Nothing in the product links or calls this file.
 *
 * Every expression below keeps the reference's IEEE operation order
 * (left-to-right evaluation, true divisions, no FMA: build with
 * -ffp-contract=off), because the pressure is recovered from the density of a
 * nearly incompressible liquid and one ulp of density is ~1/(beta*P) ulps of
 * pressure.  Each function cites the reference lines it restates.  Loops run
 * x fastest; every per-cell update of the step reads only values that no
 * other cell of the same loop writes, so the order is free (the boundary loop
 * is order-free on the grids this oracle accepts: every axis of size 1 or >= 3,
 * x >= 3 for the two-phase model).
 *
 * Pinning: the whole restatement is compared bit for bit with the compiled
 * reference (oracle/_ref) in the reference-closure mode by tests/test_oracle.py
 * and with the committed golden vectors by tests/test_golden.py.  The "corey"
 * closure has no reference counterpart; its k_r are pinned by closed forms.
 * Neither have the configurations' energy closure (energy_closure 1: the
 * two-phase energy model, constant conductivity and rock heat capacity,
 * exponential mu(T), fixed face temperatures, the z terms of energy.cpp
 * without its quirks) and the fixed bottom pressure; tests/test_oracle_config.py
 * pins them by closed forms and invariants (stationary uniform states, the
 * discrete Laplacian and advection of polynomial fields, energy conservation
 * over insulated faces, the converged Newton residual). */
#include "capaz_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>

/* Built with -fopenmp, the per-cell passes split their outer loop over
 * threads; the cells are independent (see above), so results do not depend
 * on the thread count. */
#ifdef _OPENMP
#include <omp.h>
#define OMP_PRAGMA(x) _Pragma(#x)
#define OMP_CELLS OMP_PRAGMA(omp parallel for schedule(static))
int oracle_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
#else
#define OMP_CELLS
int oracle_threads(void) { return 1; }
void oracle_set_threads(int n) { (void)n; }
#endif

#define IDX(p, i, j, k) ((size_t)(i) + (size_t)(j) * (size_t)(p)->Nx + (size_t)(k) * (size_t)(p)->Nx * (size_t)(p)->Ny)

/* ---- cell predicates ------------------------------------------------------ */

/* Interior test of assign_roS_nr (cpu.cpp:621), of the three-phase
 * assign_P_Xi/Newton (three-phase.cpp:225,288) and of energy.cpp:237,272,330,372. */
static int interior(const oracle_params* p, int i, int j, int k) {
    return (((i != 0) && (i != p->Nx - 1)) || (p->Nx < 2)) && (j != 0) && (j != p->Ny - 1) &&
           (((k != 0) && (k != p->Nz - 1)) || (p->Nz < 2));
}

/* Interior test of the two-phase Newton (two-phase.cpp:131): x has no Nx<2 escape. */
static int interior_newton2(const oracle_params* p, int i, int j, int k) {
    return (i != 0) && (i != p->Nx - 1) && (j != 0) && (j != p->Ny - 1) &&
           (((k != 0) && (k != p->Nz - 1)) || (p->Nz < 2));
}

/* Boundary tests of Border_S/Border_P: two-phase.cpp:165,195 and
 * three-phase.cpp:347,370 (the latter also energy.cpp:345). */
static int boundary(const oracle_params* p, int i, int j, int k) {
    if (p->model == 2)
        return !interior_newton2(p, i, j, k);
    return !interior(p, i, j, k);
}

/* set_boundary_basic_coordinate (cpu.cpp:107-141). */
static size_t inward(const oracle_params* p, int i, int j, int k) {
    int i1 = i, j1 = j, k1 = k;
    if (i == 0 && p->Nx > 2) i1++;
    if (i == p->Nx - 1 && p->Nx > 2) i1--;
    if (j == 0) j1++;
    if (j == p->Ny - 1) j1--;
    if (k == 0 && p->Nz > 2) k1++;
    if (k == p->Nz - 1 && p->Nz > 2) k1--;
    return IDX(p, i1, j1, k1);
}

/* ---- closures --------------------------------------------------------------- */

static double clamp01(double x) { return x < 0. ? 0. : (x > 1. ? 1. : x); }

static double corey(const oracle_params* p, double s) {
    s = clamp01(s);
    if (p->corey_n == 2.) return s * s;
    return pow(s, p->corey_n);
}

/* Three-phase effective saturations (three-phase.cpp:11-19). */
static double s_w_e3(const oracle_params* p, double S_w) {
    return (S_w - p->S_wr) / (1. - p->S_wr - p->S_nr - p->S_gr);
}
static double s_n_e3(const oracle_params* p, double S_n) {
    return (S_n - p->S_nr) / (1. - p->S_wr - p->S_nr - p->S_gr);
}

/* van Genuchten-type k_r of three-phase.cpp:146-186. */
static double vg_k_w(double A, double Swe) {
    if (Swe >= 1e-3) return pow(Swe, 0.5) * pow(1. - pow(1. - pow(Swe, A / (A - 1.)), (A - 1.) / A), 2.);
    return 0;
}
static double vg_k_g(double A, double Sge) {
    if (Sge >= 1e-3) return pow(Sge, 0.5) * pow(1. - pow(1. - Sge, A / (A - 1.)), 2. * (A - 1.) / A);
    return 0;
}
static double vg_k_n(double A, double Swe, double Sne) {
    double Sge = 1. - Swe - Sne;
    if (Sne >= 1e-3) {
        double knw = pow(1. - Swe, 0.5) * pow(1. - pow(Swe, A / (A - 1.)), 2. * (A - 1.) / A);
        double kng = pow(Sne, 0.5) * pow(1. - pow(1. - pow(Sne, A / (A - 1.)), (A - 1.) / A), 2.);
        return Sne * knw * kng / (1 - Swe) / (1 - Sge);
    }
    return 0;
}

/* Two-phase effective saturation and Brooks-Corey functions (two-phase.cpp:6-12). */
static double s_e2(const oracle_params* p, double S_n) { return (1. - S_n - p->S_wr) / (1. - p->S_wr); }

/* The "corey" closure's effective water saturation for the two-phase model:
 * (S_w - S_wr) / (1 - S_wr - S_nr) with S_w = 1 - S_n. */
static double s_e2_corey(const oracle_params* p, double S_n) {
    return (1. - S_n - p->S_wr) / (1. - p->S_wr - p->S_nr);
}

void oracle_relperm(const oracle_params* p, double S_w, double S_n, double* out) {
    if (p->model == 2) {
        if (p->closure == ORACLE_CLOSURE_COREY) {
            double se = s_e2_corey(p, S_n);
            out[0] = corey(p, se);
            out[1] = corey(p, 1. - se);
        } else {
            double lam = p->lambda, se = s_e2(p, S_n);
            out[0] = pow(se, (2. + 3. * lam) / lam);
            out[1] = (1. - se) * (1. - se) * (1 - pow(se, (2. + lam) / lam));
        }
        out[2] = 0.;
        return;
    }
    double swe = s_w_e3(p, S_w), sne = s_n_e3(p, S_n), sge = 1. - swe - sne;
    if (p->closure == ORACLE_CLOSURE_COREY) {
        out[0] = corey(p, swe);
        out[1] = corey(p, sne);
        out[2] = corey(p, sge);
    } else {
        out[0] = vg_k_w(p->lambda, swe);
        out[1] = vg_k_n(p->lambda, swe, sne);
        out[2] = vg_k_g(p->lambda, sge);
    }
}

/* ---- energy constants and closures (energy.cpp:4-143) ------------------------ */

static const double T_0 = 273.;
static const double ro_r = 2000.;
static const double lambda0_w = 0.553, lambda0_n = 0.14, lambda0_g = 0.0237, lambda0_r = 1.;
static const double lambdaA_w = 3E-3, lambdaA_n = 1E-3, lambdaA_g = 0.82;
static const double c0_w = 4.194E3, c0_n = 1.7E3, c0_g = 1E3, c0_r = 0.8E3;
static const double C_w = 1.15, C_w2 = 0.015, C_n = 3.4, C_g = 0.119, C_r = 0.75;
static const double alfa_w = 1.32E-7, alfa_n = 9.2E-7;

static double c_w(double T) { return c0_w - C_w * (T - T_0) + C_w2 * (T - T_0) * (T - T_0); }
static double c_n(double T) { return c0_n + C_n * (T - T_0); }
static double c_g(double T) { return c0_g + C_g * (T - T_0); }
static double c_r(double T) { return c0_r + C_r * (T - T_0); }
static double lambda_w(double T) { return lambda0_w * (1 - lambdaA_w * (T - T_0)); }
static double lambda_n(double T) { return lambda0_n * (1 - lambdaA_n * (T - T_0)); }
static double lambda_g(double T) { return lambda0_g * pow((T / T_0), lambdaA_g); }
static double lambda_r(double T) { (void)T; return lambda0_r; }

/* The configurations' energy closure (energy_closure 1; no reference
 * counterpart): energy.cpp's fluid properties above (H_w, H_n, H_g, c_w, c_n,
 * c_g, alfa, T_0) with a constant thermal conductivity of the medium, a
 * constant rock heat capacity and density, mu_a(T) = mu_a exp(-a (T - T_ref))
 * and prescribed face temperatures; the two-phase energy model exists only
 * with it. */
static int cfg(const oracle_params* p) { return p->energy_closure == 1 || p->model == 2; }
static double rock_ro(const oracle_params* p) { return cfg(p) ? p->ro_r : ro_r; }
static double rock_c(const oracle_params* p, double T) { return cfg(p) ? p->c_r : c_r(T); }
static double mu_of_T(const oracle_params* p, double mu0, double T) {
    return mu0 * exp(-p->mu_T_a * (T - p->mu_T_ref));
}

static double lambda_eff(const oracle_params* p, const oracle_fields* f, size_t l) {
    if (cfg(p)) return p->lambda_eff;
    double T = f->T[l];
    return f->m[l] * (f->S_w[l] * lambda_w(T) + f->S_n[l] * lambda_n(T) + f->S_g[l] * lambda_g(T)) +
           (1. - f->m[l]) * lambda_r(T);
}

static double H_w_of(const oracle_params* p, double T) {
    /* energy.cpp:114; its cubic term has the opposite sign of the integral of
     * c_w (energy.cpp:34), which the configuration closure takes instead */
    if (cfg(p)) return (p->P_atm / p->ro0_w) + (T - T_0) * (c0_w - (T - T_0) * (C_w / 2 - C_w2 * (T - T_0) / 3));
    return (p->P_atm / p->ro0_w) + (T - T_0) * (c0_w - (T - T_0) * (C_w / 2 + C_w2 * (T - T_0) / 3));
}
static double H_n_of(const oracle_params* p, double T) { return (p->P_atm / p->ro0_n) + (T - T_0) * (c0_n + C_n * (T - T_0) / 2); }
static double H_g_of(const oracle_params* p, double T) { return (p->P_atm / p->ro0_g) + (T - T_0) * (c0_g + C_g * (T - T_0) / 2); }
static double H_r_of(const oracle_params* p, double T) {
    if (cfg(p)) return (p->P_atm / p->ro_r) + (T - T_0) * p->c_r; /* the integral of a constant c_r */
    return (p->P_atm / ro_r) + (T - T_0) * (c0_r + C_r * (T - T_0) / 2);
}

static void assign_H(const oracle_params* p, oracle_fields* f, size_t l) {
    double T = f->T[l];
    f->H_w[l] = H_w_of(p, T);
    f->H_n[l] = H_n_of(p, T);
    if (p->model == 3) f->H_g[l] = H_g_of(p, T);
    f->H_r[l] = H_r_of(p, T);
}

/* ro(P, T, phase) and d_ro (energy.cpp:146-230). */
static double ro_e(const oracle_params* p, double P, double T, char ph) {
    if (ph == 'w') return p->ro0_w * (1. + (p->beta_w) * (P - p->P_atm) - alfa_w * (T - T_0));
    if (ph == 'n') return p->ro0_n * (1. + (p->beta_n) * (P - p->P_atm) - alfa_n * (T - T_0));
    return p->ro0_g * (P / p->P_atm) * (T_0 / T);
}
static double d_ro_e(const oracle_params* p, double P, double T, char ph, char var) {
    if (ph == 'w') return var == 'P' ? p->ro0_w * (p->beta_w) : (-1) * p->ro0_w * alfa_w;
    if (ph == 'n') return var == 'P' ? p->ro0_n * (p->beta_n) : (-1) * p->ro0_n * alfa_n;
    return var == 'P' ? (p->ro0_g / p->P_atm) * (T_0 / T) : p->ro0_g * (P / p->P_atm) * (-1) * (T_0 / T) / T;
}

/* ---- S_calculation (cpu.cpp:28-36, 171-182) ----------------------------------- */

void oracle_S_calculation(const oracle_params* p, oracle_fields* f) {
    size_t n = (size_t)p->Nx * p->Ny * p->Nz;
    OMP_CELLS
    for (size_t l = 0; l < n; ++l) {
        if (p->model == 3)
            f->S_g[l] = 1. - f->S_w[l] - f->S_n[l];
        else
            f->S_w[l] = 1. - f->S_n[l];
    }
}

/* ---- ro_P_Xi_calculation (cpu.cpp:3-15): assign_P_Xi + assign_ro ------------- */

static void p_xi_2(const oracle_params* p, oracle_fields* f, size_t l) {
    /* two-phase.cpp:4-23 (reference closure) */
    double k_w, k_n, P_k;
    if (p->closure == ORACLE_CLOSURE_COREY) {
        double se = s_e2_corey(p, f->S_n[l]);
        k_w = corey(p, se);
        k_n = corey(p, 1. - se);
        P_k = 0.;
    } else {
        double lam = p->lambda;
        double S_e = (1. - f->S_n[l] - p->S_wr) / (1. - p->S_wr);
        k_w = pow(S_e, (2. + 3. * lam) / lam);
        k_n = (1. - S_e) * (1. - S_e) * (1 - pow(S_e, (2. + lam) / lam));
        P_k = p->P_d * pow((1. - f->S_n[l] - p->S_wr) / (1. - p->S_wr), -1. / lam);
    }
    f->P_n[l] = f->P_w[l] + P_k;
    double mu_w = p->mu_w, mu_n = p->mu_n;
    if (p->energy) { /* the configuration's mu(T) */
        mu_w = mu_of_T(p, p->mu_w, f->T[l]);
        mu_n = mu_of_T(p, p->mu_n, f->T[l]);
    }
    f->Xi_w[l] = -1. * (f->K[l]) * k_w / mu_w;
    f->Xi_n[l] = -1. * (f->K[l]) * k_n / mu_n;
}

static void p_xi_3(const oracle_params* p, oracle_fields* f, int i, int j, int k, size_t l) {
    /* three-phase.cpp:198-240 */
    double S_w_e = s_w_e3(p, f->S_w[l]);
    double S_n_e = s_n_e3(p, f->S_n[l]);
    double S_g_e = 1. - S_w_e - S_n_e;
    double k_w, k_g, k_n;
    if (p->closure == ORACLE_CLOSURE_COREY) {
        k_w = corey(p, S_w_e);
        k_g = corey(p, S_g_e);
        k_n = corey(p, S_n_e);
    } else {
        k_w = vg_k_w(p->lambda, S_w_e);
        k_g = vg_k_g(p->lambda, S_g_e);
        k_n = vg_k_n(p->lambda, S_w_e, S_n_e);
    }
    double mu_w = p->mu_w, mu_n = p->mu_n, mu_g = p->mu_g;
    if (p->energy && cfg(p)) {
        double T = f->T[l];
        mu_w = mu_of_T(p, p->mu_w, T);
        mu_n = mu_of_T(p, p->mu_n, T);
        mu_g = mu_of_T(p, p->mu_g, T);
    } else if (p->energy) { /* three-phase.cpp:212-214 */
        double T = f->T[l];
        mu_w = 1. / (29.21 * T - 7506.64);
        mu_n = 7.256E-10 * exp(4141.9 / T);
        mu_g = 1.717E-5 * pow((T / 273.), 0.683);
    }
    f->Xi_w[l] = (-1.) * (f->K[l]) * k_w / mu_w;
    f->Xi_n[l] = (-1.) * (f->K[l]) * k_n / mu_n;
    f->Xi_g[l] = (-1.) * (f->K[l]) * k_g / mu_g;
    if (interior(p, i, j, k)) {
        /* Capillary pressures are identically zero on the CPU path
         * (three-phase.cpp:56-58, 78-80). */
        double Pk_nw = 0, Pk_gn = 0;
        f->P_n[l] = f->P_w[l] + Pk_nw;
        f->P_g[l] = f->P_w[l] + Pk_nw + Pk_gn;
    }
}

static void assign_ro(const oracle_params* p, oracle_fields* f, size_t l) {
    /* cpu.cpp:143-169 */
    if (p->model == 3) {
        if (p->energy) {
            double T0 = 273;
            f->ro_w[l] = p->ro0_w * (1. + (p->beta_w) * (f->P_w[l] - p->P_atm) - alfa_w * (f->T[l] - T0));
            f->ro_n[l] = p->ro0_n * (1. + (p->beta_n) * (f->P_n[l] - p->P_atm) - alfa_n * (f->T[l] - T0));
            f->ro_g[l] = p->ro0_g * (f->P_g[l] / p->P_atm) * (T0 / f->T[l]);
        } else {
            f->ro_w[l] = p->ro0_w * (1. + (p->beta_w) * (f->P_w[l] - p->P_atm));
            f->ro_n[l] = p->ro0_n * (1. + (p->beta_n) * (f->P_n[l] - p->P_atm));
            f->ro_g[l] = p->ro0_g * f->P_g[l] / p->P_atm;
        }
    } else if (p->energy) { /* energy.cpp:146-156's densities, two phases */
        double T0 = 273;
        f->ro_w[l] = p->ro0_w * (1. + (p->beta_w) * (f->P_w[l] - p->P_atm) - alfa_w * (f->T[l] - T0));
        f->ro_n[l] = p->ro0_n * (1. + (p->beta_n) * (f->P_n[l] - p->P_atm) - alfa_n * (f->T[l] - T0));
    } else {
        f->ro_w[l] = p->ro0_w * (1. + (p->beta_w) * (f->P_w[l] - p->P_atm));
        f->ro_n[l] = p->ro0_n * (1. + (p->beta_n) * (f->P_n[l] - p->P_atm));
    }
}

void oracle_ro_P_Xi_calculation(const oracle_params* p, oracle_fields* f) {
    OMP_CELLS
    for (int k = 0; k < p->Nz; ++k)
        for (int j = 0; j < p->Ny; ++j)
            for (int i = 0; i < p->Nx; ++i) {
                size_t l = IDX(p, i, j, k);
                if (p->model == 3)
                    p_xi_3(p, f, i, j, k, l);
                else
                    p_xi_2(p, f, l);
                assign_ro(p, f, l);
            }
}

/* ---- u_calculation (cpu.cpp:17-26, 346-477) ---------------------------------- */

static void u_axis(const oracle_params* p, const double* Xi, const double* P, const double* ro, double* u,
                   int idx, int n_axis, int n_glob, size_t l, size_t s, double h, int gravity) {
    double v;
    if (n_glob > 2) {
        double d;
        if (idx == 0)
            d = (P[l + s] - P[l]) / h;                       /* right_difference */
        else if (idx == n_axis - 1)
            d = (P[l] - P[l - s]) / h;                       /* left_difference */
        else
            d = (P[l + s] - P[l - s]) / (2. * h);            /* central_difference */
        if (gravity)
            v = Xi[l] * (d - ro[l] * (p->g_const));
        else
            v = Xi[l] * d;
    } else {
        v = 0.;
    }
    u[l] = v;
}

void oracle_u_calculation(const oracle_params* p, oracle_fields* f) {
    size_t sx = 1, sy = (size_t)p->Nx, sz = (size_t)p->Nx * p->Ny;
    OMP_CELLS
    for (int k = 0; k < p->Nz; ++k)
        for (int j = 0; j < p->Ny; ++j)
            for (int i = 0; i < p->Nx; ++i) {
                size_t l = IDX(p, i, j, k);
                u_axis(p, f->Xi_w, f->P_w, f->ro_w, f->ux_w, i, p->Nx, p->Nx, l, sx, p->hx, 0);
                u_axis(p, f->Xi_n, f->P_n, f->ro_n, f->ux_n, i, p->Nx, p->Nx, l, sx, p->hx, 0);
                u_axis(p, f->Xi_w, f->P_w, f->ro_w, f->uy_w, j, p->Ny, p->Ny, l, sy, p->hy, 1);
                u_axis(p, f->Xi_n, f->P_n, f->ro_n, f->uy_n, j, p->Ny, p->Ny, l, sy, p->hy, 1);
                u_axis(p, f->Xi_w, f->P_w, f->ro_w, f->uz_w, k, p->Nz, p->Nz, l, sz, p->hz, 0);
                u_axis(p, f->Xi_n, f->P_n, f->ro_n, f->uz_n, k, p->Nz, p->Nz, l, sz, p->hz, 0);
                if (p->model == 3) {
                    u_axis(p, f->Xi_g, f->P_g, f->ro_g, f->ux_g, i, p->Nx, p->Nx, l, sx, p->hx, 0);
                    u_axis(p, f->Xi_g, f->P_g, f->ro_g, f->uy_g, j, p->Ny, p->Ny, l, sy, p->hy, 1);
                    u_axis(p, f->Xi_g, f->P_g, f->ro_g, f->uz_g, k, p->Nz, p->Nz, l, sz, p->hz, 0);
                }
            }
}

/* ---- roS_calculation -> assign_roS_nr (cpu.cpp:52-65, 619-731) --------------- */

/* directed_difference (cpu.cpp:236-264) for one axis with stride s. */
static double directed(double x1, double x2, const double* Xi, const double* ro, size_t l, size_t s, double h) {
    return (((x2 + fabs(x2)) / 2. - (x1 - fabs(x1)) / 2.) * (-1.) * Xi[l] * ro[l] -
            (x1 + fabs(x1)) / 2. * (-1.) * Xi[l - s] * ro[l - s] +
            (x2 - fabs(x2)) / 2. * (-1.) * Xi[l + s] * ro[l + s]) / h;
}

/* Upwind flux of one phase along one axis: right/left differences of P
 * (cpu.cpp:267-314), plus g*ro_c in y (cpu.cpp:699-705). */
static double phase_flux(const oracle_params* p, const double* P, const double* Xi, const double* ro, size_t l,
                         size_t s, double h, int gravity) {
    double x2 = -1. * ((P[l + s] - P[l]) / h);
    double x1 = -1. * ((P[l] - P[l - s]) / h);
    if (gravity) {
        x2 = x2 + p->g_const * (ro[l]);
        x1 = x1 + p->g_const * (ro[l]);
    }
    return directed(x1, x2, Xi, ro, l, s, h);
}

void oracle_roS_calculation(const oracle_params* p, oracle_fields* f, double t) {
    (void)t;
    size_t sx = 1, sy = (size_t)p->Nx, sz = (size_t)p->Nx * p->Ny;
    OMP_CELLS
    for (int k = 0; k < p->Nz; ++k)
        for (int j = 0; j < p->Ny; ++j)
            for (int i = 0; i < p->Nx; ++i) {
                if (!interior(p, i, j, k)) continue;
                size_t l = IDX(p, i, j, k);
                if (!f->m[l]) continue;
                /* wells_q is identically zero for both models on the CPU path
                 * (two-phase.cpp:175-192, three-phase.cpp:529-550). */
                double q_w = 0., q_n = 0., q_g = 0.;
                if (p->model == 3) {
                    f->roS_w[l] = f->ro_w[l] * f->S_w[l];
                    f->roS_g[l] = f->ro_g[l] * (1. - f->S_w[l] - f->S_n[l]);
                } else {
                    f->roS_w[l] = f->ro_w[l] * (1. - f->S_n[l]);
                }
                f->roS_n[l] = f->ro_n[l] * f->S_n[l];
                double fx_w = 0., fy_w, fz_w = 0., fx_n = 0., fy_n, fz_n = 0., fx_g = 0., fy_g = 0., fz_g = 0.;
                if (!(p->Nz < 2)) {
                    fz_w = phase_flux(p, f->P_w, f->Xi_w, f->ro_w, l, sz, p->hz, 0);
                    fz_n = phase_flux(p, f->P_n, f->Xi_n, f->ro_n, l, sz, p->hz, 0);
                    if (p->model == 3) fz_g = phase_flux(p, f->P_g, f->Xi_g, f->ro_g, l, sz, p->hz, 0);
                }
                if (!(p->Nx < 2)) {
                    fx_w = phase_flux(p, f->P_w, f->Xi_w, f->ro_w, l, sx, p->hx, 0);
                    fx_n = phase_flux(p, f->P_n, f->Xi_n, f->ro_n, l, sx, p->hx, 0);
                    if (p->model == 3) fx_g = phase_flux(p, f->P_g, f->Xi_g, f->ro_g, l, sx, p->hx, 0);
                }
                fy_w = phase_flux(p, f->P_w, f->Xi_w, f->ro_w, l, sy, p->hy, 1);
                fy_n = phase_flux(p, f->P_n, f->Xi_n, f->ro_n, l, sy, p->hy, 1);
                double A1 = f->roS_w[l] - (p->dt / f->m[l]) * (-q_w + fx_w + fy_w + fz_w);
                double A2 = f->roS_n[l] - (p->dt / f->m[l]) * (-q_n + fx_n + fy_n + fz_n);
                f->roS_w_old[l] = f->roS_w[l];
                f->roS_n_old[l] = f->roS_n[l];
                f->roS_w[l] = A1;
                f->roS_n[l] = A2;
                if (p->model == 3) {
                    fy_g = phase_flux(p, f->P_g, f->Xi_g, f->ro_g, l, sy, p->hy, 1);
                    /* +q_g, as cpu.cpp:723 has it */
                    double A3 = f->roS_g[l] - (p->dt / f->m[l]) * (q_g + fx_g + fy_g + fz_g);
                    f->roS_g_old[l] = f->roS_g[l];
                    f->roS_g[l] = A3;
                }
            }
}

/* ---- energy: H_E_current_calculation and E_calculation (energy.cpp:235-340) -- */

void oracle_H_E_current_calculation(const oracle_params* p, oracle_fields* f) {
    size_t n = (size_t)p->Nx * p->Ny * p->Nz;
    OMP_CELLS
    for (size_t l = 0; l < n; ++l) {
        assign_H(p, f, l);
        if (p->model == 2) /* assign_E_current without the gas term */
            f->E[l] = (f->m[l] * (f->S_w[l] * (f->ro_w[l] * f->H_w[l] - f->P_w[l]) +
                                  f->S_n[l] * (f->ro_n[l] * f->H_n[l] - f->P_n[l])) +
                       (1. - f->m[l]) * (rock_ro(p) * f->H_r[l] - f->P_w[l]));
        else
            f->E[l] = (f->m[l] * (f->S_w[l] * (f->ro_w[l] * f->H_w[l] - f->P_w[l]) +
                                  f->S_n[l] * (f->ro_n[l] * f->H_n[l] - f->P_n[l]) +
                                  f->S_g[l] * (f->ro_g[l] * f->H_g[l] - f->P_g[l])) +
                       (1. - f->m[l]) * (rock_ro(p) * f->H_r[l] - f->P_w[l]));
    }
}

static double T_flow_at(const oracle_params* p, const oracle_fields* f, size_t l) {
    double T_flow = 0;
    size_t sx = 1, sy = (size_t)p->Nx, sz = (size_t)p->Nx * p->Ny;
    if (p->Nx > 2)
        T_flow += (lambda_eff(p, f, l + sx) * f->T[l + sx] - 2 * lambda_eff(p, f, l) * f->T[l] +
                   lambda_eff(p, f, l - sx) * f->T[l - sx]) / ((p->hx) * (p->hx));
    if (p->Ny > 2)
        T_flow += (lambda_eff(p, f, l + sy) * f->T[l + sy] - 2 * lambda_eff(p, f, l) * f->T[l] +
                   lambda_eff(p, f, l - sy) * f->T[l - sy]) / ((p->hy) * (p->hy));
    if (p->Nz > 2) {
        double tz = (lambda_eff(p, f, l + sz) * f->T[l + sz] - 2 * lambda_eff(p, f, l) * f->T[l] +
                     lambda_eff(p, f, l - sz) * f->T[l - sz]) / ((p->hz) * (p->hz));
        if (cfg(p))
            T_flow += tz;
        else /* '=' rather than '+=', as energy.cpp:257 has it */
            T_flow = tz;
    }
    return T_flow;
}

static double adv(const oracle_params* p, const oracle_fields* f, const double* uw, const double* un,
                  const double* ug, size_t a, size_t b) {
    if (p->model == 2) /* energy.cpp:279-306's products without the gas terms */
        return f->ro_w[a] * f->H_w[a] * uw[a] - f->ro_w[b] * f->H_w[b] * uw[b] + f->ro_n[a] * f->H_n[a] * un[a] -
               f->ro_n[b] * f->H_n[b] * un[b];
    return f->ro_w[a] * f->H_w[a] * uw[a] - f->ro_w[b] * f->H_w[b] * uw[b] + f->ro_n[a] * f->H_n[a] * un[a] -
           f->ro_n[b] * f->H_n[b] * un[b] + f->ro_g[a] * f->H_g[a] * ug[a] - f->ro_g[b] * f->H_g[b] * ug[b];
}

static double E_flow_at(const oracle_params* p, const oracle_fields* f, size_t l) {
    double E_flow = 0;
    size_t sx = 1, sy = (size_t)p->Nx, sz = (size_t)p->Nx * p->Ny;
    if (p->Nx > 2) E_flow += adv(p, f, f->ux_w, f->ux_n, f->ux_g, l + sx, l - sx) / (2. * (p->hx));
    if (p->Ny > 2) E_flow += adv(p, f, f->uy_w, f->uy_n, f->uy_g, l + sy, l - sy) / (2. * (p->hy));
    if (p->Nz > 2) {
        if (cfg(p))
            E_flow += adv(p, f, f->uz_w, f->uz_n, f->uz_g, l + sz, l - sz) / (2. * (p->hz));
        else /* the z term advects with uy, as energy.cpp:300-305 has it */
            E_flow += adv(p, f, f->uy_w, f->uy_n, f->uy_g, l + sz, l - sz) / (2. * (p->hz));
    }
    return E_flow;
}

void oracle_E_calculation(const oracle_params* p, oracle_fields* f) {
    OMP_CELLS
    for (int k = 0; k < p->Nz; ++k)
        for (int j = 0; j < p->Ny; ++j)
            for (int i = 0; i < p->Nx; ++i) {
                if (!interior(p, i, j, k)) continue;
                size_t l = IDX(p, i, j, k);
                double Q_hw = 0, Q_hr = 0;
                f->E_new[l] = f->E[l] + (p->dt) * (T_flow_at(p, f, l) + Q_hw + Q_hr - E_flow_at(p, f, l));
            }
}

/* ---- dense solve (gauss.cpp:6-104) ---------------------------------------------- */

int oracle_reverse_matrix(double* a, int n) {
    const double e = 1E-8;
    int n1 = n * 2;
    double matr[2 * 5 * 5], row[2 * 5];
    if (a == NULL || n < 2 || n > 5) return -2;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) matr[j + n1 * i] = a[j + n * i];
    for (int i = 0; i < n; i++)
        for (int j = n; j < n1; j++) matr[j + n1 * i] = (j == n + i) ? 1 : 0;
    for (int k = 0; k < n; k++) {
        double f = matr[k + n1 * k];
        if (fabs(f) < e) {
            /* Signed search that swaps with every later qualifying row
             * (gauss.cpp:39-58). */
            int found = 0;
            for (int i = k + 1; i < n; i++) {
                if (matr[k + n1 * i] > e) {
                    found = 1;
                    for (int j = 0; j < n1; j++) {
                        row[j] = matr[j + n1 * k];
                        matr[j + n1 * k] = matr[j + n1 * i];
                        matr[j + n1 * i] = row[j];
                    }
                    f = matr[k + n1 * k];
                }
            }
            if (!found) return -1; /* `a` is left unchanged */
        }
        for (int j = 0; j < n1; j++) matr[j + n1 * k] /= f;
        for (int i = 0; i < n; i++)
            if (i != k) {
                double f1 = matr[k + n1 * i];
                for (int j = 0; j < n1; j++) matr[j + n1 * i] -= (matr[j + n1 * k] * f1);
            }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) a[j + n * i] = matr[j + n + n1 * i];
    return 0;
}

void oracle_mult_matrix_vector(double* result, const double* matr, const double* vect, int n) {
    for (int i = 0; i < n; i++) {
        result[i] = 0;
        for (int j = 0; j < n; j++) result[i] += matr[j + n * i] * vect[j];
    }
}

/* ---- P_S_calculation -> Newton ----------------------------------------------------- */

static void newton2(const oracle_params* p, oracle_fields* f, size_t l) {
    /* two-phase.cpp:25-56 */
    for (int w = 1; w <= p->iterations; w++) {
        double P_k, PkS;
        if (p->closure == ORACLE_CLOSURE_COREY) {
            P_k = 0.;
            PkS = 0.;
        } else {
            double S_e = (1 - f->S_n[l] - p->S_wr) / (1 - p->S_wr);
            P_k = p->P_d * pow(S_e, (-1.) / p->lambda);
            double AAA = pow(S_e, (((-1.) / p->lambda) - 1.));
            PkS = AAA * (p->P_d) / (p->lambda * (1 - p->S_wr));
        }
        double F1 = p->ro0_w * (1. + (p->beta_w) * (f->P_w[l] - p->P_atm)) * (1. - f->S_n[l]) - f->roS_w[l];
        double F2 = p->ro0_n * (1. + (p->beta_n) * (f->P_w[l] + P_k - p->P_atm)) * f->S_n[l] - f->roS_n[l];
        double F1P = p->ro0_w * (p->beta_w) * (1 - f->S_n[l]);
        double F2P = p->ro0_n * (p->beta_n) * f->S_n[l];
        double F1S = (-1) * (p->ro0_w) * (1 + (p->beta_w) * (f->P_w[l] - p->P_atm));
        double F2S = p->ro0_n * (1 + (p->beta_n) * (f->P_w[l] + P_k - p->P_atm + (f->S_n[l] * PkS)));
        double det = F1P * F2S - F1S * F2P;
        f->P_w[l] = f->P_w[l] - (1 / det) * (F2S * F1 - F1S * F2);
        f->S_n[l] = f->S_n[l] - (1 / det) * (F1P * F2 - F2P * F1);
    }
}

static void newton3(const oracle_params* p, oracle_fields* f, size_t l) {
    /* three-phase.cpp:286-339 with P_k and their derivatives identically zero */
    double dF[9];
    for (int w = 1; w <= p->iterations; w++) {
        double Pk_nw = 0, Pk_gn = 0, PkSw = 0, PkSn = 0;
        double P = f->P_w[l], Sw = f->S_w[l], Sn = f->S_n[l];
        double Sg = 1. - Sw - Sn;
        double F1 = p->ro0_w * (1. + (p->beta_w) * (P - p->P_atm)) * Sw - f->roS_w[l];
        double F2 = p->ro0_n * (1. + (p->beta_n) * (P + Pk_nw - p->P_atm)) * Sn - f->roS_n[l];
        double F3 = p->ro0_g * (P + Pk_nw + Pk_gn) / p->P_atm * Sg - f->roS_g[l];
        dF[0] = p->ro0_w * p->beta_w * Sw;
        dF[3] = p->ro0_n * p->beta_n * Sn;
        dF[6] = p->ro0_g * Sg / p->P_atm;
        dF[1] = p->ro0_w * (1 + p->beta_w * (P - p->P_atm));
        dF[4] = p->ro0_n * (1. + (p->beta_n) * PkSw) * Sn; /* as three-phase.cpp:319 has it */
        dF[7] = (-1) * p->ro0_g * (P + Pk_nw + Pk_gn - Sg * (PkSn + PkSw)) / p->P_atm;
        dF[2] = 0;
        dF[5] = p->ro0_n * (1. + p->beta_n * (P + Pk_nw - p->P_atm));
        dF[8] = (-1) * p->ro0_g * (P + Pk_nw + Pk_gn - Sg * PkSn) / p->P_atm;
        oracle_reverse_matrix(dF, 3);
        f->P_w[l] = P - (dF[0] * F1 + dF[1] * F2 + dF[2] * F3);
        f->S_w[l] = Sw - (dF[3] * F1 + dF[4] * F2 + dF[5] * F3);
        f->S_n[l] = Sn - (dF[6] * F1 + dF[7] * F2 + dF[8] * F3);
    }
}

static void newton_energy(const oracle_params* p, oracle_fields* f, size_t l) {
    /* energy.cpp:370-463 */
    const int n = 5;
    double F[5], corr[5], dF[25];
    for (int w = 1; w <= p->iterations; w++) {
        double P = f->P_w[l], T = f->T[l], Sw = f->S_w[l], Sn = f->S_n[l], Sg = f->S_g[l], m = f->m[l];
        double rw = ro_e(p, P, T, 'w'), rn = ro_e(p, P, T, 'n'), rg = ro_e(p, P, T, 'g');
        F[0] = Sg + Sw + Sn - 1.;
        F[1] = rw * Sw - f->roS_w[l];
        F[2] = rn * Sn - f->roS_n[l];
        F[3] = rg * Sg - f->roS_g[l];
        F[4] = m * (Sw * (rw * f->H_w[l] - P) + Sn * (rn * f->H_n[l] - P) + Sg * (rg * f->H_g[l] - P)) +
               (1. - m) * (rock_ro(p) * f->H_r[l] - P) - f->E_new[l];
        dF[0] = 1.; dF[1] = 1.; dF[2] = 1.; dF[3] = 0.; dF[4] = 0.;
        dF[0 + n] = rw; dF[1 + n] = 0.; dF[2 + n] = 0.;
        dF[3 + n] = d_ro_e(p, P, T, 'w', 'P') * Sw;
        dF[4 + n] = d_ro_e(p, P, T, 'w', 'T') * Sw;
        dF[0 + n * 2] = 0.; dF[1 + n * 2] = rn; dF[2 + n * 2] = 0.;
        dF[3 + n * 2] = d_ro_e(p, P, T, 'n', 'P') * Sn;
        dF[4 + n * 2] = d_ro_e(p, P, T, 'n', 'T') * Sn;
        dF[0 + n * 3] = 0.; dF[1 + n * 3] = 0.; dF[2 + n * 3] = rg;
        dF[3 + n * 3] = d_ro_e(p, P, T, 'g', 'P') * Sg;
        dF[4 + n * 3] = d_ro_e(p, P, T, 'g', 'T') * Sg;
        dF[0 + n * 4] = m * (rw * f->H_w[l] - P);
        dF[1 + n * 4] = m * (rn * f->H_n[l] - P);
        dF[2 + n * 4] = m * (rg * f->H_g[l] - P);
        dF[3 + n * 4] = m * (Sw * (d_ro_e(p, P, T, 'w', 'P') * f->H_w[l] - 1.) +
                             Sn * (d_ro_e(p, P, T, 'n', 'P') * f->H_n[l] - 1.) +
                             Sg * (d_ro_e(p, P, T, 'g', 'P') * f->H_g[l] - 1.)) +
                        (1. - m) * (-1);
        dF[4 + n * 4] = m * (Sw * (d_ro_e(p, P, T, 'w', 'T') * f->H_w[l] + rw * c_w(T)) +
                             Sn * (d_ro_e(p, P, T, 'n', 'T') * f->H_n[l] + rn * c_n(T)) +
                             Sg * (d_ro_e(p, P, T, 'g', 'T') * f->H_g[l] + rg * c_g(T))) +
                        (1. - m) * rock_ro(p) * rock_c(p, T);
        oracle_reverse_matrix(dF, n);
        oracle_mult_matrix_vector(corr, dF, F, n);
        f->S_w[l] = Sw - corr[0];
        f->S_n[l] = Sn - corr[1];
        f->S_g[l] = Sg - corr[2];
        f->P_w[l] = P - corr[3];
        f->T[l] = T - corr[4];
        assign_H(p, f, l);
    }
    f->E[l] = f->E_new[l];
}

/* The two-phase energy model's Newton (configuration closure; no reference
 * counterpart): energy.cpp:370-463's system without the gas row and column,
 * a 4x4 system in (S_w, S_n, P_w, T) -- saturations summing to one, the two
 * mass balances and the energy balance, densities at P_w for both phases as
 * energy.cpp:388-394 has them -- inverted by reverse_matrix(dF, 4). */
static void newton_energy2(const oracle_params* p, oracle_fields* f, size_t l) {
    const int n = 4;
    double F[4], corr[4], dF[16];
    for (int w = 1; w <= p->iterations; w++) {
        double P = f->P_w[l], T = f->T[l], Sw = f->S_w[l], Sn = f->S_n[l], m = f->m[l];
        double rw = ro_e(p, P, T, 'w'), rn = ro_e(p, P, T, 'n');
        F[0] = Sw + Sn - 1.;
        F[1] = rw * Sw - f->roS_w[l];
        F[2] = rn * Sn - f->roS_n[l];
        F[3] = m * (Sw * (rw * f->H_w[l] - P) + Sn * (rn * f->H_n[l] - P)) + (1. - m) * (rock_ro(p) * f->H_r[l] - P) -
               f->E_new[l];
        dF[0] = 1.; dF[1] = 1.; dF[2] = 0.; dF[3] = 0.;
        dF[0 + n] = rw; dF[1 + n] = 0.;
        dF[2 + n] = d_ro_e(p, P, T, 'w', 'P') * Sw;
        dF[3 + n] = d_ro_e(p, P, T, 'w', 'T') * Sw;
        dF[0 + n * 2] = 0.; dF[1 + n * 2] = rn;
        dF[2 + n * 2] = d_ro_e(p, P, T, 'n', 'P') * Sn;
        dF[3 + n * 2] = d_ro_e(p, P, T, 'n', 'T') * Sn;
        dF[0 + n * 3] = m * (rw * f->H_w[l] - P);
        dF[1 + n * 3] = m * (rn * f->H_n[l] - P);
        dF[2 + n * 3] = m * (Sw * (d_ro_e(p, P, T, 'w', 'P') * f->H_w[l] - 1.) +
                             Sn * (d_ro_e(p, P, T, 'n', 'P') * f->H_n[l] - 1.)) +
                        (1. - m) * (-1);
        dF[3 + n * 3] = m * (Sw * (d_ro_e(p, P, T, 'w', 'T') * f->H_w[l] + rw * c_w(T)) +
                             Sn * (d_ro_e(p, P, T, 'n', 'T') * f->H_n[l] + rn * c_n(T))) +
                        (1. - m) * rock_ro(p) * rock_c(p, T);
        oracle_reverse_matrix(dF, n);
        oracle_mult_matrix_vector(corr, dF, F, n);
        f->S_w[l] = Sw - corr[0];
        f->S_n[l] = Sn - corr[1];
        f->P_w[l] = P - corr[2];
        f->T[l] = T - corr[3];
        assign_H(p, f, l);
    }
    f->E[l] = f->E_new[l];
}

void oracle_P_S_calculation(const oracle_params* p, oracle_fields* f) {
    OMP_CELLS
    for (int k = 0; k < p->Nz; ++k)
        for (int j = 0; j < p->Ny; ++j)
            for (int i = 0; i < p->Nx; ++i) {
                size_t l = IDX(p, i, j, k);
                if (p->model == 2) {
                    if (interior_newton2(p, i, j, k)) {
                        if (p->energy)
                            newton_energy2(p, f, l);
                        else
                            newton2(p, f, l);
                    }
                } else if (interior(p, i, j, k)) {
                    if (p->energy)
                        newton_energy(p, f, l);
                    else
                        newton3(p, f, l);
                }
            }
}

/* ---- boundary_conditions (cpu.cpp:91-104) ------------------------------------------ */

/* Border_T (energy.cpp:343-366) with the configuration's face temperatures:
 * top and bottom fixed where given (> 0), else, like the side faces,
 * insulated. */
static void border_T_cfg(const oracle_params* p, oracle_fields* f, int j, size_t l, size_t l1) {
    if (j == 0 && p->T_top > 0)
        f->T[l] = p->T_top;
    else if (j == p->Ny - 1 && p->T_bottom > 0)
        f->T[l] = p->T_bottom;
    else
        f->T[l] = f->T[l1];
}

static void border2(const oracle_params* p, oracle_fields* f, int i, int j, int k) {
    /* two-phase.cpp:163-215 and ro_eff_gdy (main.cpp:255-267) */
    size_t l1 = inward(p, i, j, k), l = IDX(p, i, j, k);
    if ((j != 0) || (p->source <= 0)) f->S_n[l] = f->S_n[l1];
    if ((j == 0) && (p->source > 0)) {
        int I = i;
        if ((I >= p->Nx / 2 - p->source) && (I <= p->Nx / 2 + p->source) && (k >= p->Nz / 2 - p->source) &&
            (k <= p->Nz / 2 + p->source))
            f->S_n[l] = p->S_n_gr;
        else
            f->S_n[l] = f->S_n[l1];
    }
    if ((j != 0) && (j != p->Ny - 1))
        f->P_w[l] = f->P_w[l1];
    else if (j == 0)
        f->P_w[l] = p->P_atm;
    else if (p->P_bottom > 0) /* the configurations' fixed bottom pressure */
        f->P_w[l] = p->P_bottom;
    else {
        double ro_g_dy = (f->ro_n[l1] * f->S_n[l1] + f->ro_w[l1] * (1 - f->S_n[l1])) * (f->m[l1]) * (p->g_const) * (p->hy);
        f->P_w[l] = f->P_w[l1] + ro_g_dy;
    }
    if (p->energy) border_T_cfg(p, f, j, l, l1);
}

static void border3(const oracle_params* p, oracle_fields* f, int i, int j, int k) {
    /* three-phase.cpp:345-427 (+ energy.cpp:343-366) */
    size_t l1 = inward(p, i, j, k), l = IDX(p, i, j, k);
    if ((j != 0) || (p->source <= 0)) {
        f->S_w[l] = f->S_w[l1];
        f->S_n[l] = f->S_n[l1];
    }
    if ((j == 0) && (p->source > 0)) {
        f->S_w[l] = p->S_w_gr;
        f->S_n[l] = p->S_n_gr;
    }
    double Pk_nw = 0, Pk_gn = 0, P1 = f->P_w[l1];
    double g = p->g_const, hy = p->hy;
    if ((j != 0) && (j != p->Ny - 1)) {
        f->P_w[l] = P1;
        f->P_n[l] = P1 + Pk_nw;
        f->P_g[l] = P1 + Pk_nw + Pk_gn;
    } else if (j == 0) {
        f->P_w[l] = (P1 - (p->ro0_w) * g * hy * (1. - (p->beta_w) * (p->P_atm))) / (1. + (p->beta_w) * (p->ro0_w) * g * hy);
        f->P_n[l] = (P1 + Pk_nw - (p->ro0_n) * g * hy * (1. - (p->beta_n) * (p->P_atm))) / (1. + (p->beta_n) * (p->ro0_n) * g * hy);
        f->P_g[l] = (P1 + Pk_nw + Pk_gn) / (1. + (p->ro0_g) * g * hy / (p->P_atm));
    } else {
        f->P_w[l] = (P1 + (p->ro0_w) * g * hy * (1. - (p->beta_w) * (p->P_atm))) / (1. - (p->beta_w) * (p->ro0_w) * g * hy);
        f->P_n[l] = (P1 + Pk_nw + (p->ro0_n) * g * hy * (1. - (p->beta_n) * (p->P_atm))) / (1. - (p->beta_n) * (p->ro0_n) * g * hy);
        f->P_g[l] = (P1 + Pk_nw + Pk_gn) / (1. - (p->ro0_g) * g * hy / (p->P_atm));
    }
    if (p->energy && cfg(p)) {
        border_T_cfg(p, f, j, l, l1);
    } else if (p->energy) {
        if (j == 0)
            f->T[l] = 400;
        else if (j == p->Ny - 1)
            f->T[l] = 273;
        else
            f->T[l] = f->T[l1];
    }
}

void oracle_boundary_conditions(const oracle_params* p, oracle_fields* f) {
    for (int i = 0; i < p->Nx; ++i)
        for (int j = 0; j < p->Ny; ++j)
            for (int k = 0; k < p->Nz; ++k) {
                if (!boundary(p, i, j, k)) continue;
                if (p->model == 2)
                    border2(p, f, i, j, k);
                else
                    border3(p, f, i, j, k);
            }
}

/* ---- time_step_function (main.cpp:93-108) --------------------------------------------- */

void oracle_time_step(const oracle_params* p, oracle_fields* f, double t) {
    oracle_S_calculation(p, f);
    oracle_ro_P_Xi_calculation(p, f);
    oracle_u_calculation(p, f);
    oracle_roS_calculation(p, f, t);
    if (p->energy) {
        oracle_H_E_current_calculation(p, f);
        oracle_E_calculation(p, f);
    }
    oracle_P_S_calculation(p, f);
    oracle_boundary_conditions(p, f);
}

void oracle_run(const oracle_params* p, oracle_fields* f, long first_step, long nsteps) {
    for (long s = first_step; s < first_step + nsteps; ++s) oracle_time_step(p, f, s * p->dt);
}
