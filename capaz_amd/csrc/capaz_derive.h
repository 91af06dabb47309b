// Host-side derivation of the device parameter block from the public
// capaz_params, and the validation of a configuration.  Constants that the
// reference evaluates inline are evaluated here with the same IEEE
// operation order (this file must be compiled without FMA contraction), so
// the kernels read the same bits the reference computes.
#pragma once

#include <string>

#include "../../include/capaz.h"
#include "capaz_params.h"

namespace capaz {

// Empty string if the configuration is supported, else the reason.
inline std::string validate(const capaz_params& u) {
    if (u.struct_size != sizeof(capaz_params)) return "capaz_params.struct_size does not match this library";
    if (u.model != 2 && u.model != 3) return "model must be CAPAZ_TWO_PHASE (2) or CAPAZ_THREE_PHASE (3)";
    if (u.energy != 0 && u.energy != 1) return "energy must be 0 or 1";
    if (u.energy_closure != 0 && u.energy_closure != 1)
        return "energy_closure must be CAPAZ_ENERGY_REFERENCE or CAPAZ_ENERGY_CONFIG";
    if (u.energy && u.model != 3 && u.energy_closure != 1)
        return "the two-phase energy equation needs energy_closure CAPAZ_ENERGY_CONFIG (energy.cpp is three-phase)";
    if (u.energy && u.energy_closure == 1 && !(u.lambda_eff > 0 && u.c_r > 0 && u.ro_r > 0))
        return "CAPAZ_ENERGY_CONFIG needs lambda_eff, c_r and ro_r > 0";
    if (u.closure != 0 && u.closure != 1) return "closure must be CAPAZ_CLOSURE_REFERENCE or CAPAZ_CLOSURE_COREY";
    if (u.Ny < 3) return "Ny must be >= 3";
    if (!(u.Nx == 1 || u.Nx >= 3)) return "Nx must be 1 or >= 3";
    if (u.model == 2 && u.Nx < 3) return "the two-phase model needs Nx >= 3 (two-phase.cpp:131,165)";
    if (!(u.Nz == 1 || u.Nz >= 3)) return "Nz must be 1 or >= 3";
    if ((double)u.Nx * (double)u.Ny * (double)u.Nz > 4.0e9) return "grid too large";
    if (u.newton_iterations < 0) return "newton_iterations must be >= 0";
    if (!(u.hx > 0 && u.hy > 0 && u.hz > 0)) return "hx, hy, hz must be > 0";
    if (!(u.dt > 0)) return "dt must be > 0";
    if (!(u.P_atm > 0)) return "P_atm must be > 0";
    if (!(u.mu_w > 0 && u.mu_n > 0)) return "mu_w and mu_n must be > 0";
    if (u.model == 3 && (!u.energy || u.energy_closure == 1) && !(u.mu_g > 0)) return "mu_g must be > 0";
    if (u.P_bottom < 0) return "P_bottom must be >= 0";
    if (u.closure == 1 && !(u.corey_n > 0)) return "corey_n must be > 0";
    if (u.model == 2 && u.closure == 0 && !(u.lambda != 0)) return "lambda must be non-zero";
    if (u.model == 3 && u.closure == 0 && !(u.lambda != 0 && u.lambda != 1)) return "lambda must not be 0 or 1";
    return "";
}

inline DevParams derive(const capaz_params& u) {
    DevParams p = {};
    p.model = u.model;
    p.energy = u.energy;
    p.closure = u.closure;
    p.Nx = u.Nx;
    p.Ny = u.Ny;
    p.Nz = u.Nz;
    p.source = u.source;
    p.iterations = u.newton_iterations;
    p.sy = u.Nx;
    p.sz = (int64_t)u.Nx * u.Ny;
    p.n = p.sz * u.Nz;
    p.hx = u.hx;
    p.hy = u.hy;
    p.hz = u.hz;
    p.dt = u.dt;
    p.P_atm = u.P_atm;
    p.g = u.g_const;
    p.mu_w = u.mu_w;
    p.mu_n = u.mu_n;
    p.mu_g = u.mu_g;
    p.ro0_w = u.ro0_w;
    p.ro0_n = u.ro0_n;
    p.ro0_g = u.ro0_g;
    p.beta_w = u.beta_w;
    p.beta_n = u.beta_n;
    p.lambda = u.lambda;
    p.S_wr = u.S_wr;
    p.S_nr = u.S_nr;
    p.S_gr = u.S_gr;
    p.P_d = u.P_d;
    p.S_w_gr = u.S_w_gr;
    p.S_n_gr = u.S_n_gr;
    p.corey_n = u.corey_n;

    p.r_hx = 1. / u.hx;
    p.r_hy = 1. / u.hy;
    p.r_hz = 1. / u.hz;
    p.two_hx = 2. * u.hx;
    p.two_hy = 2. * u.hy;
    p.two_hz = 2. * u.hz;
    p.r_two_hx = 1. / p.two_hx;
    p.r_two_hy = 1. / p.two_hy;
    p.r_two_hz = 1. / p.two_hz;
    p.hx2 = u.hx * u.hx;
    p.hy2 = u.hy * u.hy;
    p.hz2 = u.hz * u.hz;
    p.r_hx2 = 1. / p.hx2;
    p.r_hy2 = 1. / p.hy2;
    p.r_hz2 = 1. / p.hz2;
    p.r_mu_w = 1. / u.mu_w;
    p.r_mu_n = 1. / u.mu_n;
    p.r_mu_g = u.mu_g != 0 ? 1. / u.mu_g : 0.;
    p.r_P_atm = 1. / u.P_atm;

    // three-phase.cpp:13,18
    p.den3 = 1. - u.S_wr - u.S_nr - u.S_gr;
    p.r_den3 = 1. / p.den3;
    // two-phase.cpp:6; the Corey closure's (1 - S_wr - S_nr)
    p.den2 = 1. - u.S_wr;
    p.r_den2 = 1. / p.den2;
    p.den2c = 1. - u.S_wr - u.S_nr;
    p.r_den2c = 1. / p.den2c;
    // two-phase.cpp:11-13, 36-44
    const double lam = u.lambda;
    p.bc_ew = (2. + 3. * lam) / lam;
    p.bc_en = (2. + lam) / lam;
    p.bc_epk = -1. / lam;
    p.bc_eaaa = ((-1.) / lam) - 1.;
    p.bc_pks_den = lam * (1 - u.S_wr);
    p.r_bc_pks_den = 1. / p.bc_pks_den;
    // three-phase.cpp:153-181
    const double A = u.lambda;
    p.vg_a = A / (A - 1.);
    p.vg_b = (A - 1.) / A;
    p.vg_c = 2. * (A - 1.) / A;
    // three-phase.cpp:315-316
    p.c_wb = u.ro0_w * u.beta_w;
    p.c_nb = u.ro0_n * u.beta_n;
    // three-phase.cpp:402-421
    const double g = u.g_const, hy = u.hy;
    p.bp_w_num = (u.ro0_w) * g * hy * (1. - (u.beta_w) * (u.P_atm));
    p.bp_w_top_den = 1. + (u.beta_w) * (u.ro0_w) * g * hy;
    p.bp_w_bot_den = 1. - (u.beta_w) * (u.ro0_w) * g * hy;
    p.bp_n_num = (u.ro0_n) * g * hy * (1. - (u.beta_n) * (u.P_atm));
    p.bp_n_top_den = 1. + (u.beta_n) * (u.ro0_n) * g * hy;
    p.bp_n_bot_den = 1. - (u.beta_n) * (u.ro0_n) * g * hy;
    p.bp_g_top_den = 1. + (u.ro0_g) * g * hy / (u.P_atm);
    p.bp_g_bot_den = 1. - (u.ro0_g) * g * hy / (u.P_atm);
    p.r_bp_w_top_den = 1. / p.bp_w_top_den;
    p.r_bp_w_bot_den = 1. / p.bp_w_bot_den;
    p.r_bp_n_top_den = 1. / p.bp_n_top_den;
    p.r_bp_n_bot_den = 1. / p.bp_n_bot_den;
    p.r_bp_g_top_den = 1. / p.bp_g_top_den;
    p.r_bp_g_bot_den = 1. / p.bp_g_bot_den;
    // energy.cpp:114-129
    p.H0_w = u.P_atm / u.ro0_w;
    p.H0_n = u.P_atm / u.ro0_n;
    p.H0_g = u.ro0_g != 0 ? u.P_atm / u.ro0_g : 0.;
    p.ecfg = u.energy && u.energy_closure == 1;
    p.ro_r = p.ecfg ? u.ro_r : 2000.;
    p.H0_r = u.P_atm / p.ro_r;
    p.P_bottom = u.P_bottom;
    p.lam_eff = u.lambda_eff;
    p.c_r = u.c_r;
    p.T_top = u.T_top;
    p.T_bottom = u.T_bottom;
    p.mu_T_a = u.mu_T_a;
    p.mu_T_ref = u.mu_T_ref;
    return p;
}

}  // namespace capaz
