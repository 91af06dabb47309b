// C ABI of the backend (include/capaz.h): context, device memory of every
// field, the entry-point sequence and the fused step.
//
// State model.  Every field of the model has one device array, the current
// one, on which the entry-point kernels work in place exactly as the
// reference's loops work on its host arrays.  The primary state (P_w, S_n,
// and for three phases S_w, P_n, P_g) has a second array: a fused step reads
// the current arrays and writes the second ones, then the two swap.  A fused
// step does not store the intermediate fields (S_g / S_w, Xi, ro, u, roS, the
// interior P_n / P_g); they are marked stale and recomputed from the previous
// state -- which the second arrays still hold -- before anything reads them,
// so every field always reads as after the same number of reference steps.
// The three-phase energy model's step (six passes) works in place like the
// sequence; the two-phase energy model's fused step ping-pongs P_w, S_n and T
// and re-runs the step's entry points from the previous state when a caller
// needs the intermediates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <utility>

#include "capaz_ctx.h"

using capaz::DevFields;
using capaz::DevParams;



namespace {
thread_local std::string g_create_error;
}  // namespace

namespace capaz::detail {

int fail(capaz_ctx* c, int code, const std::string& msg) {
    if (c)
        c->err = msg;
    else
        g_create_error = msg;
    return code;
}

int hip_fail(capaz_ctx* c, hipError_t e, const char* what) {
    return fail(c, e == hipErrorOutOfMemory ? CAPAZ_ERR_NOMEM : CAPAZ_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace capaz::detail

namespace {

using capaz::detail::fail;
using capaz::detail::hip_fail;

#define CAPAZ_HIP(c, call)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
    } while (0)

bool field_exists(const capaz_params& u, int f) {
    if (f < 0 || f >= capaz::F_COUNT) return false;
    if (f == capaz::F_H_g) return u.energy == 1 && u.model == 3;
    if (f >= capaz::F_T) return u.energy == 1;
    if (f >= capaz::F_S_g) return u.model == 3;
    return true;
}

// Fields whose value carries over from step to step (the rest are recomputed
// from these at every step).
bool is_primary(const capaz_params& u, int f) {
    using namespace capaz;
    if (f == F_P_w || f == F_S_n) return true;
    if (u.model == 3 && (f == F_S_w || f == F_P_n || f == F_P_g)) return true;
    if (u.energy && f == F_T) return true;
    return false;
}

DevFields fields_of(capaz_ctx* c) {
    DevFields F;
    for (int f = 0; f < capaz::F_COUNT; ++f) F.f[f] = c->cur[f];
    return F;
}

int run_entry(capaz_ctx* c, capaz::Entry e, DevFields F) {
    CAPAZ_HIP(c, capaz::launch_entry(e, c->p, F, c->stream));
    return CAPAZ_OK;
}

// Recomputes the intermediate fields of the last fused step from the
// previous primary state (held in the second arrays).
int materialize(capaz_ctx* c) {
    if (!c->stale) return CAPAZ_OK;
    DevFields F = fields_of(c);
    for (int f = 0; f < capaz::F_COUNT; ++f)
        if (c->primary[f]) F.f[f] = c->alt[f];
    int rc;
    if (c->user.energy && c->user.model == 2) {
        // two-phase energy model: the Newton's S_w, H and E are intermediates
        // too, so the whole step runs again from the previous state; its new
        // state lands in the second arrays with the bits the current ones hold
        using capaz::Entry;
        for (Entry e : {Entry::S, Entry::P_Xi_ro, Entry::U, Entry::RoS, Entry::HE, Entry::E})
            if ((rc = run_entry(c, e, F))) return rc;
        CAPAZ_HIP(c, capaz::launch_newton_energy(c->p, F, c->tuning.energy_newton ? c->d_redo : nullptr, c->stream));
        if ((rc = run_entry(c, Entry::Boundary, F))) return rc;
        c->stale = false;
        return CAPAZ_OK;
    }
    if ((rc = run_entry(c, capaz::Entry::S, F))) return rc;
    if ((rc = run_entry(c, capaz::Entry::P_Xi_ro, F))) return rc;
    if ((rc = run_entry(c, capaz::Entry::U, F))) return rc;
    if ((rc = run_entry(c, capaz::Entry::RoS, F))) return rc;
    if (c->user.model == 3) {
        CAPAZ_HIP(c, capaz::launch_copy_interior(c->p, c->alt[capaz::F_P_n], c->cur[capaz::F_P_n], c->stream));
        CAPAZ_HIP(c, capaz::launch_copy_interior(c->p, c->alt[capaz::F_P_g], c->cur[capaz::F_P_g], c->stream));
    }
    c->stale = false;
    return CAPAZ_OK;
}

int check_ctx(capaz_ctx* c) {
    if (!c) return CAPAZ_ERR_INVALID;
    CAPAZ_HIP(c, hipSetDevice(c->user.device));
    return CAPAZ_OK;
}

int check_field(capaz_ctx* c, int field, size_t n) {
    if (!field_exists(c->user, field)) return fail(c, CAPAZ_ERR_INVALID, "field does not exist for this model");
    if (n != (size_t)c->p.n) return fail(c, CAPAZ_ERR_INVALID, "n must equal Nx*Ny*Nz");
    return CAPAZ_OK;
}

constexpr size_t kEventPool = 512;

// Waits for the recorded event pairs and adds their elapsed times to the totals.
int drain_events(capaz_ctx* c) {
    if (!c->ev_used) return CAPAZ_OK;
    CAPAZ_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->ev_used; ++i) {
        float ms = 0.f;
        CAPAZ_HIP(c, hipEventElapsedTime(&ms, c->ev[i].begin, c->ev[i].end));
        c->timed_ms += ms;
    }
    c->timed_launches += (long)c->ev_used;
    c->ev_used = 0;
    return CAPAZ_OK;
}

// The next free event pair of the pool, or NULL when timing is off.
int next_events(capaz_ctx* c, const capaz::KernelEvents** out) {
    *out = nullptr;
    if (!c->timing) return CAPAZ_OK;
    if (c->ev.empty()) {
        c->ev.resize(kEventPool, capaz::KernelEvents{nullptr, nullptr});
        for (auto& e : c->ev) {
            CAPAZ_HIP(c, hipEventCreate(&e.begin));
            CAPAZ_HIP(c, hipEventCreate(&e.end));
        }
    }
    int rc;
    if (c->ev_used == c->ev.size() && (rc = drain_events(c))) return rc;
    *out = &c->ev[c->ev_used++];
    return CAPAZ_OK;
}

// The non-finite sentinel over the primary state (P_w, S_n; S_w for three
// phases; T with the energy equation): *first = the smallest cell index
// holding a NaN or an infinity in one of them, -1 if none.
int find_nonfinite(capaz_ctx* c, int64_t* first) {
    using namespace capaz;
    const double* f[4] = {c->cur[F_P_w], c->cur[F_S_n], c->user.model == 3 ? c->cur[F_S_w] : nullptr,
                          c->user.energy ? c->cur[F_T] : nullptr};
    CAPAZ_HIP(c, launch_find_nonfinite(c->p, f, c->d_first, c->stream));
    unsigned long long h = 0;
    CAPAZ_HIP(c, hipMemcpyAsync(&h, c->d_first, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    CAPAZ_HIP(c, hipStreamSynchronize(c->stream));
    *first = h == ~0ull ? -1 : (int64_t)h;
    return CAPAZ_OK;
}

// CAPAZ_ERR_NONFINITE with a message naming cell l, its (i, j, k) and the
// state fields that are not finite there (after step `step`, if >= 0).
int nonfinite_error(capaz_ctx* c, int64_t l, long step) {
    using namespace capaz;
    const int ids[4] = {F_P_w, F_S_n, F_S_w, F_T};
    const char* names[4] = {"P_w", "S_n", "S_w", "T"};
    std::string which;
    for (int q = 0; q < 4; ++q) {
        if ((q == 2 && c->user.model != 3) || (q == 3 && !c->user.energy)) continue;
        double v = 0.;
        CAPAZ_HIP(c, hipMemcpyAsync(&v, c->cur[ids[q]] + l, sizeof(v), hipMemcpyDeviceToHost, c->stream));
        CAPAZ_HIP(c, hipStreamSynchronize(c->stream));
        if (!std::isfinite(v)) which += std::string(which.empty() ? "" : ", ") + names[q] + " = " + std::to_string(v);
    }
    const int64_t i = l % c->p.Nx, j = (l / c->p.Nx) % c->p.Ny, k = l / ((int64_t)c->p.Nx * c->p.Ny);
    std::string msg = "non-finite state at cell " + std::to_string(l) + " (i " + std::to_string(i) + ", j " +
                      std::to_string(j) + ", k " + std::to_string(k) + "): " + which;
    if (step >= 0) msg += " after step " + std::to_string(step);
    return fail(c, CAPAZ_ERR_NONFINITE, msg);
}

// The sentinel of capaz_step after step number `step` (every check_every-th).
int check_after_step(capaz_ctx* c, long step) {
    const int every = c->tuning.check_every;
    if (every <= 0 || step % every != 0) return CAPAZ_OK;
    int64_t first = -1;
    int rc;
    if ((rc = find_nonfinite(c, &first))) return rc;
    return first < 0 ? CAPAZ_OK : nonfinite_error(c, first, step);
}

int scan_m(capaz_ctx* c) {
    if (c->m_checked) return CAPAZ_OK;
    CAPAZ_HIP(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
    CAPAZ_HIP(c, capaz::launch_m_zero(c->p, c->cur[capaz::F_m], c->d_flag, c->stream));
    CAPAZ_HIP(c, capaz::launch_dtm(c->p, c->cur[capaz::F_m], c->d_dtm, c->stream));
    int h = 0;
    CAPAZ_HIP(c, hipMemcpyAsync(&h, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAPAZ_HIP(c, hipStreamSynchronize(c->stream));
    c->m_has_zero = h != 0;
    c->m_checked = true;
    return CAPAZ_OK;
}

}  // namespace

namespace capaz::detail {
int materialize(capaz_ctx* c) { return ::materialize(c); }
int scan_m(capaz_ctx* c) { return ::scan_m(c); }
int check_ctx(capaz_ctx* c) { return ::check_ctx(c); }
}  // namespace capaz::detail

extern "C" {

void capaz_params_init(capaz_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->struct_size = sizeof(capaz_params);
    p->model = CAPAZ_THREE_PHASE;
    p->newton_iterations = 1;
    p->corey_n = 2.;
}

int capaz_abi_version(void) { return CAPAZ_ABI_VERSION; }

int capaz_create(const capaz_params* u, capaz_ctx** out) {
    if (!u || !out) return fail(nullptr, CAPAZ_ERR_INVALID, "NULL argument");
    *out = nullptr;
    std::string why = capaz::validate(*u);
    if (!why.empty()) return fail(nullptr, CAPAZ_ERR_INVALID, why);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(nullptr, CAPAZ_ERR_DEVICE, "no HIP device available");
    if (u->device < 0 || u->device >= ndev) return fail(nullptr, CAPAZ_ERR_INVALID, "device ordinal out of range");
    capaz_ctx* c = new (std::nothrow) capaz_ctx;
    if (!c) return fail(nullptr, CAPAZ_ERR_NOMEM, "host allocation failed");
    c->user = *u;
    c->p = capaz::derive(*u);
    auto bail = [&](hipError_t err, const char* what) {
        g_create_error = std::string(what) + ": " + hipGetErrorString(err);
        capaz_destroy(c);
        return err == hipErrorOutOfMemory ? CAPAZ_ERR_NOMEM : CAPAZ_ERR_DEVICE;
    };
    if ((e = hipSetDevice(u->device)) != hipSuccess) return bail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "hipStreamCreate");
    if ((e = hipMalloc(&c->d_flag, sizeof(int))) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMalloc(&c->d_first, sizeof(unsigned long long))) != hipSuccess) return bail(e, "hipMalloc");
    const size_t bytes = (size_t)c->p.n * sizeof(double);
    if ((e = hipMalloc(&c->d_dtm, bytes)) != hipSuccess) return bail(e, "hipMalloc");
    if (u->energy && (e = hipMalloc(&c->d_lam, bytes)) != hipSuccess) return bail(e, "hipMalloc");
    if (u->energy && (e = hipMalloc(&c->d_redo, ((size_t)c->p.n + 15) / 16 * 16)) != hipSuccess)
        return bail(e, "hipMalloc");
    for (int f = 0; f < capaz::F_COUNT; ++f) {
        if (!field_exists(*u, f)) continue;
        if ((e = hipMalloc(&c->cur[f], bytes)) != hipSuccess) return bail(e, "hipMalloc");
        if ((e = hipMemsetAsync(c->cur[f], 0, bytes, c->stream)) != hipSuccess) return bail(e, "hipMemset");
        // ping-pong: the fused steps (the three-phase energy step works in place)
        c->primary[f] = is_primary(*u, f) && !(u->energy && u->model == 3);
        if (c->primary[f]) {
            if ((e = hipMalloc(&c->alt[f], bytes)) != hipSuccess) return bail(e, "hipMalloc");
            if ((e = hipMemsetAsync(c->alt[f], 0, bytes, c->stream)) != hipSuccess) return bail(e, "hipMemset");
        }
    }
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    *out = c;
    return CAPAZ_OK;
}

int capaz_destroy(capaz_ctx* c) {
    if (!c) return CAPAZ_OK;
    (void)hipSetDevice(c->user.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int f = 0; f < capaz::F_COUNT; ++f) {
        if (c->cur[f]) (void)hipFree(c->cur[f]);
        if (c->alt[f]) (void)hipFree(c->alt[f]);
    }
    if (c->d_flag) (void)hipFree(c->d_flag);
    if (c->d_first) (void)hipFree(c->d_first);
    if (c->d_dtm) (void)hipFree(c->d_dtm);
    if (c->d_lam) (void)hipFree(c->d_lam);
    if (c->d_redo) (void)hipFree(c->d_redo);
    if (c->adj_stream) (void)hipStreamSynchronize(c->adj_stream);
    if (c->adj_mem) (void)hipFree(c->adj_mem);
    for (hipEvent_t e : c->adj_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->adj_stream) (void)hipStreamDestroy(c->adj_stream);
    for (auto& e : c->ev) {
        if (e.begin) (void)hipEventDestroy(e.begin);
        if (e.end) (void)hipEventDestroy(e.end);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CAPAZ_OK;
}

const char* capaz_last_error(const capaz_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

size_t capaz_cells(const capaz_ctx* c) { return c ? (size_t)c->p.n : 0; }

int capaz_has_field(const capaz_ctx* c, int field) { return c && field_exists(c->user, field) ? 1 : 0; }

int capaz_set_field(capaz_ctx* c, int field, const double* host, size_t n) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if ((rc = check_field(c, field, n))) return rc;
    if (!host) return fail(c, CAPAZ_ERR_INVALID, "NULL host pointer");
    if ((rc = materialize(c))) return rc;
    CAPAZ_HIP(c, hipMemcpyAsync(c->cur[field], host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CAPAZ_HIP(c, hipStreamSynchronize(c->stream));
    if (field == capaz::F_m) c->m_checked = false;
    return CAPAZ_OK;
}

int capaz_set_field_device(capaz_ctx* c, int field, const double* dsrc, size_t n) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if ((rc = check_field(c, field, n))) return rc;
    if (!dsrc) return fail(c, CAPAZ_ERR_INVALID, "NULL device pointer");
    if ((rc = materialize(c))) return rc;
    CAPAZ_HIP(c, hipMemcpyAsync(c->cur[field], dsrc, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (field == capaz::F_m) c->m_checked = false;
    return CAPAZ_OK;
}

int capaz_get_field(capaz_ctx* c, int field, double* host, size_t n) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if ((rc = check_field(c, field, n))) return rc;
    if (!host) return fail(c, CAPAZ_ERR_INVALID, "NULL host pointer");
    if ((rc = materialize(c))) return rc;
    CAPAZ_HIP(c, hipMemcpyAsync(host, c->cur[field], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CAPAZ_HIP(c, hipStreamSynchronize(c->stream));
    return CAPAZ_OK;
}

int capaz_field_device_ptr(capaz_ctx* c, int field, double** dptr) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (!field_exists(c->user, field)) return fail(c, CAPAZ_ERR_INVALID, "field does not exist for this model");
    if (!dptr) return fail(c, CAPAZ_ERR_INVALID, "NULL dptr");
    if ((rc = materialize(c))) return rc;
    *dptr = c->cur[field];
    if (field == capaz::F_m) c->m_checked = false;  // the caller may write through the pointer
    return CAPAZ_OK;
}

#define CAPAZ_ENTRY(name, entry, guard)                                          \
    int name(capaz_ctx* c) {                                                     \
        int rc;                                                                  \
        if ((rc = check_ctx(c))) return rc;                                      \
        if (!(guard)) return fail(c, CAPAZ_ERR_STATE, #name ": not in this model"); \
        if ((rc = materialize(c))) return rc;                                    \
        return run_entry(c, capaz::Entry::entry, fields_of(c));                  \
    }

CAPAZ_ENTRY(capaz_S, S, true)
CAPAZ_ENTRY(capaz_P_Xi_ro, P_Xi_ro, true)
CAPAZ_ENTRY(capaz_u, U, true)
CAPAZ_ENTRY(capaz_energy_HE, HE, c->user.energy)

int capaz_energy_E(capaz_ctx* c) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (!c->user.energy) return fail(c, CAPAZ_ERR_STATE, "capaz_energy_E: not in this model");
    if ((rc = materialize(c))) return rc;
    if (c->user.model == 2) return run_entry(c, capaz::Entry::E, fields_of(c));
    // on the z tiles, lambda_eff evaluated once per staged cell
    CAPAZ_HIP(c, capaz::launch_energy_e_tiles(c->p, fields_of(c), nullptr, c->stream));
    return CAPAZ_OK;
}
int capaz_newton(capaz_ctx* c) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if ((rc = materialize(c))) return rc;
    if (c->user.energy) {
        CAPAZ_HIP(c, capaz::launch_newton_energy(c->p, fields_of(c), c->tuning.energy_newton ? c->d_redo : nullptr,
                                                 c->stream));
        return CAPAZ_OK;
    }
    return run_entry(c, capaz::Entry::Newton, fields_of(c));
}
CAPAZ_ENTRY(capaz_boundary, Boundary, true)

int capaz_roS(capaz_ctx* c, double t) {
    (void)t;  // assign_roS_nr does not read t (cpu.cpp:619-731)
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if ((rc = materialize(c))) return rc;
    if (c->user.model == 3) {
        // three phases: on the z tiles unless some interior cell has m == 0
        if ((rc = scan_m(c))) return rc;
        if (!c->m_has_zero) {
            CAPAZ_HIP(c, capaz::launch_roS_stored(c->p, fields_of(c), c->d_dtm, c->stream));
            return CAPAZ_OK;
        }
    }
    return run_entry(c, capaz::Entry::RoS, fields_of(c));
}

int capaz_step_sequence(capaz_ctx* c, double t) {
    int rc;
    if ((rc = capaz_S(c))) return rc;
    if ((rc = capaz_P_Xi_ro(c))) return rc;
    if ((rc = capaz_u(c))) return rc;
    if ((rc = capaz_roS(c, t))) return rc;
    if (c->user.energy) {
        if ((rc = capaz_energy_HE(c))) return rc;
        if ((rc = capaz_energy_E(c))) return rc;
    }
    if ((rc = capaz_newton(c))) return rc;
    return capaz_boundary(c);
}

int capaz_step(capaz_ctx* c, long first_step, long nsteps) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (nsteps < 0) return fail(c, CAPAZ_ERR_INVALID, "nsteps must be >= 0");
    if ((rc = scan_m(c))) return rc;
    using namespace capaz;
    if (c->halo_pack && nsteps > 0 && (c->user.energy || c->m_has_zero))
        return fail(c, CAPAZ_ERR_INVALID,
                    "halo pack: the fused step only (no energy equation, no interior cell with m == 0)");
    if (c->user.energy && c->user.model == 3 && !c->m_has_zero && c->tuning.energy_kernel >= 1) {
        // the energy model's six passes, in place like the sequence
        if ((rc = materialize(c))) return rc;
        for (long s = 0; s < nsteps; ++s) {
            const KernelEvents* ev;
            if ((rc = next_events(c, &ev))) return rc;
            CAPAZ_HIP(c, launch_energy_step(c->p, fields_of(c), c->d_lam, c->d_dtm,
                                            c->tuning.energy_newton ? c->d_redo : nullptr, c->stream, ev));
            if ((rc = check_after_step(c, first_step + s))) return rc;
        }
        return CAPAZ_OK;
    }
    if (c->user.energy && c->user.model == 2 && !c->m_has_zero && c->tuning.energy_kernel >= 1) {
        // the two-phase energy model's fused step (ping-pong P_w, S_n, T)
        for (long s = 0; s < nsteps; ++s) {
            const double* cur[5] = {c->cur[F_P_w], c->cur[F_S_n], c->cur[F_T], c->cur[F_m], c->cur[F_K]};
            double* next[3] = {c->alt[F_P_w], c->alt[F_S_n], c->alt[F_T]};
            const KernelEvents* ev;
            if ((rc = next_events(c, &ev))) return rc;
            CAPAZ_HIP(c, launch_energy2_step(c->p, cur, next, c->tuning.kchunk, c->stream, ev));
            for (int f = 0; f < F_COUNT; ++f)
                if (c->primary[f]) std::swap(c->cur[f], c->alt[f]);
            c->stale = true;
            if ((rc = check_after_step(c, first_step + s))) return rc;
        }
        return CAPAZ_OK;
    }
    if (c->user.energy || c->m_has_zero) {
        // Interior cells with m == 0 keep the reference's stale-roS rule,
        // which only the sequence reproduces.
        for (long s = first_step; s < first_step + nsteps; ++s) {
            if ((rc = capaz_step_sequence(c, (double)s * c->user.dt))) return rc;
            if ((rc = check_after_step(c, s))) return rc;
        }
        return CAPAZ_OK;
    }
    for (long s = 0; s < nsteps; ++s) {
        StepIO io;
        io.Pw = c->cur[F_P_w];
        io.Sw = c->cur[F_S_w];
        io.Sn = c->cur[F_S_n];
        io.Pn = c->cur[F_P_n];
        io.Pg = c->cur[F_P_g];
        io.m = c->cur[F_m];
        io.K = c->cur[F_K];
        io.dtm = c->d_dtm;
        io.Pw1 = c->alt[F_P_w];
        io.Sw1 = c->alt[F_S_w];
        io.Sn1 = c->alt[F_S_n];
        io.Pn1 = c->alt[F_P_n];
        io.Pg1 = c->alt[F_P_g];
        io.pack = c->halo_pack;
        const KernelEvents* ev;
        if ((rc = next_events(c, &ev))) return rc;
        CAPAZ_HIP(c, launch_fused(c->p, io, c->tuning, c->stream, ev));
        for (int f = 0; f < F_COUNT; ++f)
            if (c->primary[f]) std::swap(c->cur[f], c->alt[f]);
        c->stale = true;
        if ((rc = check_after_step(c, first_step + s))) return rc;
    }
    return CAPAZ_OK;
}

int capaz_tune(capaz_ctx* c, const char* key, int value) {
    if (!c || !key) return CAPAZ_ERR_INVALID;
    const std::string k(key);
    if (k == "check_every" && value >= 0)
        c->tuning.check_every = value;
    else if (k == "fused_kernel" && (value >= 0 && value <= 2))
        c->tuning.fused_kernel = value;
    else if (k == "tile_y" && (value == 0 || value == 4 || value == 8))
        c->tuning.tile_y = value;
    else if (k == "kchunk" && value >= 0 && value <= 1 << 20)
        c->tuning.kchunk = value;
    else if (k == "adjoint_kernel" && value >= 0 && value <= 2)
        c->tuning.adjoint_kernel = value;
    else if (k == "adjoint_streams" && (value == 1 || value == 2))
        c->tuning.adjoint_streams = value;
    else if (k == "energy_kernel" && (value == 0 || value == 1))
        c->tuning.energy_kernel = value;
    else if (k == "energy_newton" && (value == 0 || value == 1))
        c->tuning.energy_newton = value;
    else if (k == "timing" && (value == 0 || value == 1)) {
        int rc;
        if ((rc = check_ctx(c)) || (rc = drain_events(c))) return rc;
        c->timing = value == 1;
        c->timed_ms = 0.;
        c->timed_launches = 0;
    } else
        return fail(c, CAPAZ_ERR_INVALID, "capaz_tune: unknown key or value out of range");
    return CAPAZ_OK;
}

int capaz_get_tuning(const capaz_ctx* c, const char* key, int* value) {
    if (!c || !key || !value) return CAPAZ_ERR_INVALID;
    const std::string k(key);
    if (k == "check_every") *value = c->tuning.check_every;
    else if (k == "fused_kernel") *value = c->tuning.fused_kernel;
    else if (k == "tile_y") *value = c->tuning.tile_y;
    else if (k == "kchunk") *value = c->tuning.kchunk;
    else if (k == "adjoint_kernel") *value = c->tuning.adjoint_kernel;
    else if (k == "adjoint_streams") *value = c->tuning.adjoint_streams;
    else if (k == "energy_kernel") *value = c->tuning.energy_kernel;
    else if (k == "energy_newton") *value = c->tuning.energy_newton;
    else if (k == "timing") *value = c->timing ? 1 : 0;
    else if (k == "adjoint_segment") *value = (int)std::min<long>(c->adj_segment, 0x7fffffff);
    else return CAPAZ_ERR_INVALID;
    return CAPAZ_OK;
}

int capaz_kernel_time(capaz_ctx* c, double* total_ms, long* launches) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (!total_ms || !launches) return fail(c, CAPAZ_ERR_INVALID, "NULL argument");
    if ((rc = drain_events(c))) return rc;
    *total_ms = c->timed_ms;
    *launches = c->timed_launches;
    return CAPAZ_OK;
}

int capaz_check_finite(capaz_ctx* c, int64_t* first_bad) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (!first_bad) return fail(c, CAPAZ_ERR_INVALID, "NULL first_bad");
    *first_bad = -1;
    if ((rc = find_nonfinite(c, first_bad))) return rc;
    return *first_bad < 0 ? CAPAZ_OK : nonfinite_error(c, *first_bad, -1);
}

int capaz_sync(capaz_ctx* c) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    CAPAZ_HIP(c, hipStreamSynchronize(c->stream));
    return CAPAZ_OK;
}

int capaz_stream(capaz_ctx* c, void** s) {
    if (!c || !s) return CAPAZ_ERR_INVALID;
    *s = (void*)c->stream;
    return CAPAZ_OK;
}

// Where the interior cells of a state field's current array are stale after
// fused steps: P_n and P_g of the three-phase model, whose interior values
// after a step are the previous P_w (P_Xi_ro, three-phase.cpp:230-231, before
// the Newton) -- the second P_w array, after the step's swap -- while the
// fused step stores only their boundary values.  NULL where the array is
// current.
static const double* stale_inner(capaz_ctx* c, int field) {
    if (!c->stale || c->user.model != 3 || !c->primary[field]) return nullptr;
    return field == capaz::F_P_n || field == capaz::F_P_g ? c->alt[capaz::F_P_w] : nullptr;
}

static int face_call(capaz_ctx* c, int field, int axis, int side, double* buf, int save) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (!field_exists(c->user, field)) return fail(c, CAPAZ_ERR_INVALID, "field does not exist for this model");
    if (axis < 0 || axis > 2 || (side != 0 && side != 1) || !buf)
        return fail(c, CAPAZ_ERR_INVALID, "axis must be 0..2, side 0 or 1, buffer non-NULL");
    int n = axis == 0 ? c->p.Nx : (axis == 1 ? c->p.Ny : c->p.Nz);
    if (n < 2) return fail(c, CAPAZ_ERR_INVALID, "axis has fewer than 2 planes");
    // a field of the ping-pong state is current as it is (its ghost planes do
    // not enter the last step's intermediates, which come from the previous
    // state); the others first need the last step's intermediates
    if (!c->primary[field] && (rc = materialize(c))) return rc;
    // load_*: second / second-to-last plane; save_*: first / last plane
    int pos = save ? (side == 0 ? 0 : n - 1) : (side == 0 ? 1 : n - 2);
    CAPAZ_HIP(c, capaz::launch_face(c->p, c->cur[field], save ? nullptr : stale_inner(c, field), buf, axis, pos,
                                    save, c->stream));
    if (save && field == capaz::F_m) c->m_checked = false;
    return CAPAZ_OK;
}

int capaz_load_face(capaz_ctx* c, int field, int axis, int side, double* dbuf) {
    return face_call(c, field, axis, side, dbuf, 0);
}

int capaz_save_face(capaz_ctx* c, int field, int axis, int side, const double* dbuf) {
    return face_call(c, field, axis, side, const_cast<double*>(dbuf), 1);
}

static int faces_call(capaz_ctx* c, const int* fields, int nfields, int axis, double* buf, int save) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (!fields || !buf || nfields < 1 || nfields > capaz::kMaxFaceFields)
        return fail(c, CAPAZ_ERR_INVALID, "fields and buffer non-NULL, 1 to 8 fields");
    if (axis < 0 || axis > 2) return fail(c, CAPAZ_ERR_INVALID, "axis must be 0..2");
    const int n = axis == 0 ? c->p.Nx : (axis == 1 ? c->p.Ny : c->p.Nz);
    if (n < 2) return fail(c, CAPAZ_ERR_INVALID, "axis has fewer than 2 planes");
    double* ptr[capaz::kMaxFaceFields];
    bool all_primary = true;
    for (int q = 0; q < nfields; ++q) {
        if (!field_exists(c->user, fields[q])) return fail(c, CAPAZ_ERR_INVALID, "field does not exist for this model");
        all_primary = all_primary && c->primary[fields[q]];
    }
    if (!all_primary && (rc = materialize(c))) return rc;
    const double* inner[capaz::kMaxFaceFields];
    for (int q = 0; q < nfields; ++q) {
        ptr[q] = c->cur[fields[q]];
        inner[q] = save ? nullptr : stale_inner(c, fields[q]);
        if (save && fields[q] == capaz::F_m) c->m_checked = false;
    }
    CAPAZ_HIP(c, capaz::launch_faces(c->p, ptr, inner, nfields, buf, axis, save, c->stream));
    return CAPAZ_OK;
}

int capaz_set_halo_pack(capaz_ctx* c, double* dbuf) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (dbuf) {
        if (c->user.energy) return fail(c, CAPAZ_ERR_INVALID, "halo pack: models without the energy equation only");
        if (c->p.Nz < 3) return fail(c, CAPAZ_ERR_INVALID, "halo pack: Nz must be at least 3");
        if (reinterpret_cast<uintptr_t>(dbuf) % alignof(double))
            return fail(c, CAPAZ_ERR_INVALID, "halo pack: buffer not aligned to 8 bytes");
    }
    c->halo_pack = dbuf;
    return CAPAZ_OK;
}

int capaz_ghost_planes(capaz_ctx* c, double** dst, int nmax, int* n) {
    int rc;
    if ((rc = check_ctx(c))) return rc;
    if (!dst || !n) return fail(c, CAPAZ_ERR_INVALID, "ghost planes: NULL argument");
    if (c->user.energy) return fail(c, CAPAZ_ERR_INVALID, "ghost planes: models without the energy equation only");
    if (c->p.Nz < 3) return fail(c, CAPAZ_ERR_INVALID, "ghost planes: Nz must be at least 3");
    using namespace capaz;
    static const int three[] = {F_P_w, F_S_w, F_S_n, F_P_n, F_P_g}, two[] = {F_P_w, F_S_n};
    const int* f = c->user.model == 3 ? three : two;
    const int nf = c->user.model == 3 ? 5 : 2;
    if (nmax < 2 * nf) return fail(c, CAPAZ_ERR_INVALID, "ghost planes: dst has room for fewer than 2 x fields");
    for (int q = 0; q < nf; ++q) {
        dst[2 * q] = c->cur[f[q]];
        dst[2 * q + 1] = c->cur[f[q]] + (int64_t)(c->p.Nz - 1) * c->p.sz;
    }
    *n = 2 * nf;
    return CAPAZ_OK;
}

int capaz_load_faces(capaz_ctx* c, const int* fields, int nfields, int axis, double* dbuf) {
    return faces_call(c, fields, nfields, axis, dbuf, 0);
}

int capaz_save_faces(capaz_ctx* c, const int* fields, int nfields, int axis, const double* dbuf) {
    return faces_call(c, fields, nfields, axis, const_cast<double*>(dbuf), 1);
}

}  // extern "C"
