// Test-only harness around the compiled reference CPU path (the oracle's
// anchor).  It is compiled together with the reference's own sources
// (main.cpp, cpu.cpp, gauss.cpp, the model .cpp, no_communication.cpp,
// shared_test.cpp and, for the energy build, energy.cpp) by oracle/Makefile
// into oracle/_ref/libcapaz_ref_<model>.so.  Nothing in the product links it;
// only tests, the golden-vector script and the CPU-baseline leg of bench.py
// load it, to compare against.
//
// The harness owns one single-domain reference context: a `consts` block set
// from defines.ini keys (the key names and their meaning follow
// main.cpp:1327-1626) and the `ptr_Arrays` host arrays allocated by the
// reference's own host_memory_allocation (main.cpp:365-426).  Every entry
// point the time step uses is callable on its own, so each of the backend's
// kernels can be checked against the exact reference function it replaces.
#include "defines.h"

#include <cstring>
#include <string>

namespace {

struct RefCtx {
    consts def;
    ptr_Arrays a;
    bool allocated;
};

size_t cells(const RefCtx* c) {
    return (size_t)c->def.locNx * (size_t)c->def.locNy * (size_t)c->def.locNz;
}

// Maps a defines.ini key onto the consts field it sets (main.cpp:1327-1626).
int set_key(consts* d, const char* k, double v) {
    const std::string key(k);
    if (key == "HX") d->hx = v;
    else if (key == "HY") d->hy = v;
    else if (key == "HZ") d->hz = v;
    else if (key == "TAU") d->tau = v;
    else if (key == "DT") d->dt = v;
    else if (key == "C_W") d->c_w = v;
    else if (key == "C_N") d->c_n = v;
    else if (key == "L") d->l = v;
    else if (key == "BETA_W") d->beta_w = v;
    else if (key == "BETA_N") d->beta_n = v;
    else if (key == "RO_W") d->ro0_w = v;
    else if (key == "RO_N") d->ro0_n = v;
    else if (key == "MU_W") d->mu_w = v;
    else if (key == "MU_N") d->mu_n = v;
    else if (key == "G_CONST") d->g_const = v;
    else if (key == "P_ATM") d->P_atm = v;
    else if (key == "LAMBDA_0") d->lambda[0] = v;
    else if (key == "LAMBDA_1") d->lambda[1] = v;
    else if (key == "M_0") d->porosity[0] = v;
    else if (key == "M_1") d->porosity[1] = v;
    else if (key == "S_WR_0") d->S_wr[0] = v;
    else if (key == "S_WR_1") d->S_wr[1] = v;
    else if (key == "Q") d->Q = v;
    else if (key == "K_0") d->K[0] = v;
    else if (key == "K_1") d->K[1] = v;
    else if (key == "S_N_GR") d->S_n_gr = v;
    else if (key == "SOURCE") d->source = (int)v;
    else if (key == "ITERATIONS") d->newton_iterations = (int)v;
    else if (key == "TIMEX") d->timeX = v;
    else if (key == "SAVE_PLOTS") d->save_plots = (int)v;
    else if (key == "PRINT_SCREEN") d->print_screen = (int)v;
    else if (key == "NX") d->Nx = (int)v;
    else if (key == "NY") d->Ny = (int)v;
    else if (key == "NZ") d->Nz = (int)v;
#ifdef TWO_PHASE
    else if (key == "P_D_0") d->P_d[0] = v;
    else if (key == "P_D_1") d->P_d[1] = v;
#endif
#ifdef THREE_PHASE
    else if (key == "C_G") d->c_g = v;
    else if (key == "BETA_G") d->beta_g = v;
    else if (key == "RO_G") d->ro0_g = v;
    else if (key == "MU_G") d->mu_g = v;
    else if (key == "P_D_NW_0") d->P_d_nw[0] = v;
    else if (key == "P_D_NW_1") d->P_d_nw[1] = v;
    else if (key == "P_D_GN_0") d->P_d_gn[0] = v;
    else if (key == "P_D_GN_1") d->P_d_gn[1] = v;
    else if (key == "S_W_GR") d->S_w_gr = v;
    else if (key == "S_W_INIT") d->S_w_init = v;
    else if (key == "S_N_INIT") d->S_n_init = v;
    else if (key == "S_NR_0") d->S_nr[0] = v;
    else if (key == "S_NR_1") d->S_nr[1] = v;
    else if (key == "S_GR_0") d->S_gr[0] = v;
    else if (key == "S_GR_1") d->S_gr[1] = v;
#endif
    else return -1;
    return 0;
}

double** field_slot(ptr_Arrays* a, const char* n) {
    const std::string s(n);
#define CAPAZ_REF_FIELD(x) if (s == #x) return &a->x;
    CAPAZ_REF_FIELD(S_n) CAPAZ_REF_FIELD(S_w) CAPAZ_REF_FIELD(P_w) CAPAZ_REF_FIELD(P_n)
    CAPAZ_REF_FIELD(ro_w) CAPAZ_REF_FIELD(ro_n)
    CAPAZ_REF_FIELD(ux_w) CAPAZ_REF_FIELD(uy_w) CAPAZ_REF_FIELD(uz_w)
    CAPAZ_REF_FIELD(ux_n) CAPAZ_REF_FIELD(uy_n) CAPAZ_REF_FIELD(uz_n)
    CAPAZ_REF_FIELD(Xi_w) CAPAZ_REF_FIELD(Xi_n)
    CAPAZ_REF_FIELD(roS_w) CAPAZ_REF_FIELD(roS_w_old) CAPAZ_REF_FIELD(roS_n) CAPAZ_REF_FIELD(roS_n_old)
    CAPAZ_REF_FIELD(m) CAPAZ_REF_FIELD(K)
#ifdef THREE_PHASE
    CAPAZ_REF_FIELD(S_g) CAPAZ_REF_FIELD(P_g) CAPAZ_REF_FIELD(ro_g)
    CAPAZ_REF_FIELD(ux_g) CAPAZ_REF_FIELD(uy_g) CAPAZ_REF_FIELD(uz_g)
    CAPAZ_REF_FIELD(Xi_g) CAPAZ_REF_FIELD(roS_g) CAPAZ_REF_FIELD(roS_g_old)
#endif
#ifdef ENERGY
    CAPAZ_REF_FIELD(T) CAPAZ_REF_FIELD(H_w) CAPAZ_REF_FIELD(H_n) CAPAZ_REF_FIELD(H_g)
    CAPAZ_REF_FIELD(H_r) CAPAZ_REF_FIELD(E) CAPAZ_REF_FIELD(E_new)
#endif
#undef CAPAZ_REF_FIELD
    return nullptr;
}

const char* const kFieldNames[] = {
    "S_n", "S_w", "P_w", "P_n", "ro_w", "ro_n", "ux_w", "uy_w", "uz_w", "ux_n", "uy_n", "uz_n",
    "Xi_w", "Xi_n", "roS_w", "roS_w_old", "roS_n", "roS_n_old", "m", "K",
#ifdef THREE_PHASE
    "S_g", "P_g", "ro_g", "ux_g", "uy_g", "uz_g", "Xi_g", "roS_g", "roS_g_old",
#endif
#ifdef ENERGY
    "T", "H_w", "H_n", "H_g", "H_r", "E", "E_new",
#endif
};

}  // namespace

extern "C" {

const char* capaz_ref_model(void) {
#if defined(THREE_PHASE) && defined(ENERGY)
    return "3phe";
#elif defined(THREE_PHASE)
    return "3ph";
#else
    return "2ph";
#endif
}

int capaz_ref_field_count(void) { return (int)(sizeof(kFieldNames) / sizeof(kFieldNames[0])); }

const char* capaz_ref_field_name(int i) {
    if (i < 0 || i >= capaz_ref_field_count()) return nullptr;
    return kFieldNames[i];
}

void* capaz_ref_create(void) {
    RefCtx* c = new RefCtx;
    std::memset(&c->def, 0, sizeof(c->def));
    std::memset(&c->a, 0, sizeof(c->a));
    c->allocated = false;
    c->def.upscale_l = 1;  // read_defines sets these after parsing (main.cpp:1628-1629)
    c->def.upscale_t = 1;
    return c;
}

int capaz_ref_set(void* ctx, const char* key, double value) {
    RefCtx* c = static_cast<RefCtx*>(ctx);
    if (c->allocated) return -2;
    return set_key(&c->def, key, value);
}

// Same set-up order as initialization() (main.cpp:296-312) on one process,
// then zero-filled host arrays so that fields a caller does not set are
// deterministic.  If run_data_initialization is non-zero the model's own
// data_initialization fills the state, as on a run without save/save.dat.
int capaz_ref_init(void* ctx, int run_data_initialization) {
    RefCtx* c = static_cast<RefCtx*>(ctx);
    if (c->allocated) return -2;
    communication_initialization(0, nullptr, &c->def);
    sizes_initialization(&c->def);
    blocks_initialization(&c->def);
    global_to_local_vars(&c->def);
    host_memory_allocation(&c->a, c->def);
    c->allocated = true;
    for (int i = 0; i < capaz_ref_field_count(); ++i) {
        double** slot = field_slot(&c->a, kFieldNames[i]);
        std::memset(*slot, 0, cells(c) * sizeof(double));
    }
    load_permeability(c->a.K, c->def);
    if (run_data_initialization) {
        long int tc = 0;
        data_initialization(c->a, &tc, c->def);
    }
    return 0;
}

long capaz_ref_cells(void* ctx) { return (long)cells(static_cast<RefCtx*>(ctx)); }

double* capaz_ref_field(void* ctx, const char* name) {
    RefCtx* c = static_cast<RefCtx*>(ctx);
    if (!c->allocated) return nullptr;
    double** slot = field_slot(&c->a, name);
    return slot ? *slot : nullptr;
}

// One reference entry point, by the name of the defines.h function it calls.
int capaz_ref_call(void* ctx, const char* entry, double t) {
    RefCtx* c = static_cast<RefCtx*>(ctx);
    if (!c->allocated) return -2;
    const std::string e(entry);
    if (e == "S_calculation") S_calculation(c->a, c->a, c->def);
    else if (e == "ro_P_Xi_calculation") ro_P_Xi_calculation(c->a, c->a, c->def);
    else if (e == "u_calculation") u_calculation(c->a, c->a, c->def);
    else if (e == "roS_calculation") roS_calculation(c->a, c->a, t, c->def);
    else if (e == "P_S_calculation") P_S_calculation(c->a, c->a, c->def);
    else if (e == "boundary_conditions") boundary_conditions(c->a, c->a, c->def);
#ifdef ENERGY
    else if (e == "H_E_current_calculation") H_E_current_calculation(c->a, c->a, c->def);
    else if (e == "E_calculation") E_calculation(c->a, c->a, c->def);
#endif
    else if (e == "time_step_function") time_step_function(c->a, c->a, nullptr, c->def, t);
    else return -1;
    return 0;
}

// Runs `nsteps` calls of time_step_function with t = step * dt, the way the
// driver loop in main.cpp:41-58 does, starting at step number `first_step`.
int capaz_ref_run(void* ctx, long first_step, long nsteps) {
    RefCtx* c = static_cast<RefCtx*>(ctx);
    if (!c->allocated) return -2;
    for (long s = first_step; s < first_step + nsteps; ++s)
        time_step_function(c->a, c->a, nullptr, c->def, s * c->def.dt);
    return 0;
}

int capaz_ref_reverse_matrix(double* a, int n) { return reverse_matrix(a, n); }

void capaz_ref_mult_matrix_vector(double* r, double* m, double* v, int n) { mult_matrix_vector(r, m, v, n); }

void capaz_ref_destroy(void* ctx) {
    RefCtx* c = static_cast<RefCtx*>(ctx);
    if (c->allocated) host_memory_free(c->a);
    delete c;
}

}  // extern "C"
