// Device-side parameter block of one domain, passed to every kernel by value
// (it lands in the kernel-argument segment and is read with scalar loads).
//
// Besides the user parameters it carries constants derived on the host with
// the same IEEE operation order the reference uses when it evaluates them
// inline, so the device reads bit-identical values, plus correctly rounded
// reciprocals RN(1/b) of every constant divisor for exact_div().
#pragma once

#include <stdint.h>

namespace capaz {

struct DevParams {
    int model, energy, closure;
    int Nx, Ny, Nz;
    int source, iterations;
    int64_t sy, sz, n;  // strides of j and k, number of cells
    double hx, hy, hz, dt;
    double P_atm, g;
    double mu_w, mu_n, mu_g;
    double ro0_w, ro0_n, ro0_g;
    double beta_w, beta_n;
    double lambda, S_wr, S_nr, S_gr, P_d;
    double S_w_gr, S_n_gr, corey_n;

    // reciprocals of constant divisors: r_x = RN(1/x)
    double r_hx, r_hy, r_hz;
    double two_hx, two_hy, two_hz, r_two_hx, r_two_hy, r_two_hz;  // 2.*h (central_difference)
    double hx2, hy2, hz2, r_hx2, r_hy2, r_hz2;                    // h*h (energy.cpp:246)
    double r_mu_w, r_mu_n, r_mu_g;
    double r_P_atm;

    // three-phase effective-saturation denominator 1. - S_wr - S_nr - S_gr
    double den3, r_den3;
    // two-phase: 1. - S_wr (Brooks-Corey) and 1. - S_wr - S_nr (Corey)
    double den2, r_den2, den2c, r_den2c;
    // two-phase Brooks-Corey exponents and P_k' denominator (two-phase.cpp:11-13,36-44)
    double bc_ew, bc_en, bc_epk, bc_eaaa, bc_pks_den, r_bc_pks_den;
    // van Genuchten exponents A/(A-1.), (A-1.)/A, 2.*(A-1.)/A (three-phase.cpp:153-181)
    double vg_a, vg_b, vg_c;
    // Newton constants: ro0_w*beta_w, ro0_n*beta_n (three-phase.cpp:315-316)
    double c_wb, c_nb;
    // three-phase Border_P top/bottom constants (three-phase.cpp:402-421)
    double bp_w_num, bp_w_top_den, bp_w_bot_den, r_bp_w_top_den, r_bp_w_bot_den;
    double bp_n_num, bp_n_top_den, bp_n_bot_den, r_bp_n_top_den, r_bp_n_bot_den;
    double bp_g_top_den, bp_g_bot_den, r_bp_g_top_den, r_bp_g_bot_den;
    // energy: P_atm/ro0_* and P_atm/ro_r (energy.cpp:114-129)
    double H0_w, H0_n, H0_g, H0_r;

    // configuration closures (no reference counterpart): the two-phase fixed
    // bottom pressure (> 0) and the energy closure (ecfg = 1): constant
    // conductivity, rock heat capacity and density, face temperatures
    // (<= 0: insulated) and mu(T) = mu0 exp(-mu_T_a (T - mu_T_ref)).  In the
    // reference energy closure c_r and lam_eff are unused and ro_r = 2000.
    double P_bottom;
    int ecfg;
    double lam_eff, c_r, ro_r, T_top, T_bottom, mu_T_a, mu_T_ref;
};

}  // namespace capaz
