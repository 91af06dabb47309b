/* The oracle: a CPU restatement, in plain C, of capaz's explicit filtration
 * step (cpu.cpp + Two-phase/two-phase.cpp + Three-phase/three-phase.cpp +
 * gauss.cpp + energy.cpp).  Test infrastructure only: tests, the
 * golden-vector script, smoke() and the CPU-baseline leg of bench.py call it
 * to compare against; the product never links it.
 *
 * Layout: one double array per field, x fastest, local = i + j*Nx + k*Nx*Ny
 * (cpu.cpp:145).  Fields that a model does not have are NULL. */
#ifndef CAPAZ_ORACLE_H
#define CAPAZ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_CLOSURE_REFERENCE = 0, ORACLE_CLOSURE_COREY = 1 };

typedef struct oracle_params {
    int model;      /* 2: water/NAPL, 3: water/NAPL/gas */
    int energy;     /* 1: energy equation (energy.cpp for model 3; model 2
                       with energy_closure 1 only) */
    int closure;    /* ORACLE_CLOSURE_* */
    int Nx, Ny, Nz;
    int source;     /* SOURCE */
    int iterations; /* ITERATIONS */
    double hx, hy, hz, dt;
    double P_atm, g_const;
    double mu_w, mu_n, mu_g;
    double ro0_w, ro0_n, ro0_g;
    double beta_w, beta_n;
    double lambda;  /* LAMBDA_0 */
    double S_wr, S_nr, S_gr; /* S_WR_0, S_NR_0, S_GR_0 */
    double P_d;     /* P_D_0 (two-phase) */
    double S_w_gr, S_n_gr;
    double corey_n; /* Corey exponent of the "corey" closure */
    /* Boundary and energy closures of the benchmark configurations (no
     * reference counterpart; zero values select the reference's rules). */
    double P_bottom;    /* two phases: > 0 a fixed bottom-face pressure instead of
                           the hydrostatic rule of two-phase.cpp:100-107 */
    int energy_closure; /* 0: energy.cpp's constants, closures and its z-term
                           quirks (three phases only); 1: the configuration's
                           closure (below), for two or three phases */
    double lambda_eff;  /* 1: thermal conductivity of the medium, W/(m K) */
    double c_r, ro_r;   /* 1: rock heat capacity J/(kg K), rock density kg/m^3 */
    double T_top, T_bottom; /* 1: face temperatures; <= 0 insulated (copy inward) */
    double mu_T_a, mu_T_ref; /* 1: mu_a(T) = mu_a * exp(-mu_T_a * (T - mu_T_ref)) */
} oracle_params;

typedef struct oracle_fields {
    double *S_n, *S_w, *P_w, *P_n, *ro_w, *ro_n;
    double *ux_w, *uy_w, *uz_w, *ux_n, *uy_n, *uz_n;
    double *Xi_w, *Xi_n, *roS_w, *roS_w_old, *roS_n, *roS_n_old, *m, *K;
    double *S_g, *P_g, *ro_g, *ux_g, *uy_g, *uz_g, *Xi_g, *roS_g, *roS_g_old;
    double *T, *H_w, *H_n, *H_g, *H_r, *E, *E_new;
} oracle_fields;

/* Entry points, one per defines.h function of the step (cpu.cpp:3-104). */
void oracle_S_calculation(const oracle_params* p, oracle_fields* f);
void oracle_ro_P_Xi_calculation(const oracle_params* p, oracle_fields* f);
void oracle_u_calculation(const oracle_params* p, oracle_fields* f);
void oracle_roS_calculation(const oracle_params* p, oracle_fields* f, double t);
void oracle_H_E_current_calculation(const oracle_params* p, oracle_fields* f);
void oracle_E_calculation(const oracle_params* p, oracle_fields* f);
void oracle_P_S_calculation(const oracle_params* p, oracle_fields* f);
void oracle_boundary_conditions(const oracle_params* p, oracle_fields* f);
/* time_step_function (main.cpp:93-108) on one domain. */
void oracle_time_step(const oracle_params* p, oracle_fields* f, double t);
/* nsteps calls with t = s*dt, s = first_step .. first_step+nsteps-1. */
void oracle_run(const oracle_params* p, oracle_fields* f, long first_step, long nsteps);

/* Threads the per-cell passes use (1 unless built with OpenMP). */
int oracle_threads(void);
void oracle_set_threads(int n);

/* gauss.cpp:6-92 and :94-104. */
int oracle_reverse_matrix(double* a, int n);
void oracle_mult_matrix_vector(double* result, const double* matr, const double* vect, int n);

/* Relative permeabilities of the configured closure at one saturation
 * state, for closed-form checks: out[0..2] = k_w, k_n, k_g (k_g = 0 for the
 * two-phase model, whose only saturation input is S_n). */
void oracle_relperm(const oracle_params* p, double S_w, double S_n, double* out);

#ifdef __cplusplus
}
#endif
#endif
