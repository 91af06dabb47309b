// C ABI of the history-matching objective and its discrete adjoint
// (include/capaz.h: capaz_objective, capaz_adjoint).
//
// capaz_adjoint runs the forward steps from the context's state, keeping a
// checkpoint of the primary state every C steps and the observed saturations
// at the sampled steps, then sweeps the segments backwards: each segment is
// recomputed from its checkpoint with the fused kernel (the same bits as the
// first pass), recording each cell's flux divergence, and its steps are
// reversed with the two adjoint passes.  The last segment is recorded by the
// forward sweep itself, so with C = nsteps nothing is recomputed.
// Device memory: (nsteps/C + C + 1) states of five fields, 3C recorded
// fields and 18 fields of adjoint work, held by the context between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "capaz_adjoint_launch.h"
#include "capaz_ctx.h"

namespace {

using capaz::detail::fail;
using capaz::detail::hip_fail;

#define ADJ_HIP(c, call)                                      \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return hip_fail((c), e_, #call); \
    } while (0)

constexpr int kState = 5;  // P_w, S_w, S_n, P_n, P_g
const int kStateField[kState] = {capaz::F_P_w, capaz::F_S_w, capaz::F_S_n, capaz::F_P_n, capaz::F_P_g};

struct Obs {
    std::vector<int> sample_of;  // step -> sample index or -1
    int64_t ncells = 0;
    int nsamples = 0;
};

int check_model(capaz_ctx* c) {
    if (c->user.energy) return fail(c, CAPAZ_ERR_INVALID, "adjoint: models without the energy equation only");
    if (c->user.model == 2 && c->user.closure != CAPAZ_CLOSURE_COREY)
        return fail(c, CAPAZ_ERR_INVALID, "adjoint: two phases with the Corey closure only");
    if (c->user.model == 2 && c->tuning.adjoint_kernel != 2)
        return fail(c, CAPAZ_ERR_INVALID, "adjoint: two phases with adjoint_kernel 2 only");
    if (c->p.iterations > capaz::kMaxAdjIter)
        return fail(c, CAPAZ_ERR_INVALID, "adjoint: at most 8 Newton iterations");
    if (c->p.Ny > 65535 || c->p.Nz > 65535) return fail(c, CAPAZ_ERR_INVALID, "adjoint: Ny, Nz <= 65535");
    return CAPAZ_OK;
}

int check_obs(capaz_ctx* c, long nsteps, const capaz_observations* o, Obs* out) {
    if (nsteps < 1) return fail(c, CAPAZ_ERR_INVALID, "nsteps must be >= 1");
    if (!o) return fail(c, CAPAZ_ERR_INVALID, "NULL observations");
    if (o->nsamples < 0) return fail(c, CAPAZ_ERR_INVALID, "observations: nsamples < 0");
    if ((o->ncells && !o->cells) || (o->nsamples && (!o->steps || (o->ncells && (!o->S_w || !o->S_n)))))
        return fail(c, CAPAZ_ERR_INVALID, "observations: NULL array");
    out->sample_of.assign((size_t)nsteps + 1, -1);
    long prev = 0;
    for (int s = 0; s < o->nsamples; ++s) {
        if (o->steps[s] <= prev || o->steps[s] > nsteps)
            return fail(c, CAPAZ_ERR_INVALID, "observations: steps must increase within 1..nsteps");
        prev = o->steps[s];
        out->sample_of[(size_t)o->steps[s]] = s;
    }
    std::vector<int64_t> sorted(o->cells, o->cells + o->ncells);
    std::sort(sorted.begin(), sorted.end());
    for (size_t q = 0; q < sorted.size(); ++q) {
        if (sorted[q] < 0 || sorted[q] >= c->p.n) return fail(c, CAPAZ_ERR_INVALID, "observations: cell out of range");
        if (q && sorted[q] == sorted[q - 1]) return fail(c, CAPAZ_ERR_INVALID, "observations: repeated cell");
    }
    out->ncells = (int64_t)o->ncells;
    out->nsamples = o->nsamples;
    return CAPAZ_OK;
}

// Device workspace of the context, grown on demand (fields of `stride` doubles).
int workspace(capaz_ctx* c, size_t doubles, double** out) {
    const size_t bytes = doubles * sizeof(double);
    if (c->adj_bytes < bytes) {
        ADJ_HIP(c, hipStreamSynchronize(c->stream));
        if (c->adj_stream) ADJ_HIP(c, hipStreamSynchronize(c->adj_stream));
        if (c->adj_mem) (void)hipFree(c->adj_mem);
        c->adj_mem = nullptr;
        c->adj_bytes = 0;
        ADJ_HIP(c, hipMalloc(&c->adj_mem, bytes));
        c->adj_bytes = bytes;
    }
    *out = c->adj_mem;
    return CAPAZ_OK;
}

// One fused step from the five arrays at `from` to those at `to` (stride
// apart) on stream s; with div != NULL it also records the flux divergence
// (3 fields).
int fused_between(capaz_ctx* c, hipStream_t s, const double* from, double* to, size_t stride, double* div = nullptr) {
    capaz::StepIO io;
    io.Pw = from;
    io.Sw = from + stride;
    io.Sn = from + 2 * stride;
    io.Pn = from + 3 * stride;
    io.Pg = from + 4 * stride;
    io.m = c->cur[capaz::F_m];
    io.K = c->cur[capaz::F_K];
    io.dtm = c->d_dtm;
    io.Pw1 = to;
    io.Sw1 = to + stride;
    io.Sn1 = to + 2 * stride;
    io.Pn1 = to + 3 * stride;
    io.Pg1 = to + 4 * stride;
    if (div)
        ADJ_HIP(c, capaz::launch_fused_record(c->p, io, c->tuning.kchunk, div, (int64_t)stride, s));
    else
        ADJ_HIP(c, capaz::launch_fused(c->p, io, c->tuning, s));
    return CAPAZ_OK;
}

// The state fields the model keeps between steps: all five for three phases,
// P_w and S_n for two (their slots keep the five-field layout).
bool state_field(const capaz_ctx* c, int f) { return c->p.model == 3 || f == 0 || f == 2; }

int copy_state(capaz_ctx* c, double* dst, size_t stride) {
    const size_t bytes = (size_t)c->p.n * sizeof(double);
    for (int f = 0; f < kState; ++f)
        if (state_field(c, f))
            ADJ_HIP(c, hipMemcpyAsync(dst + f * stride, c->cur[kStateField[f]], bytes, hipMemcpyDeviceToDevice,
                                      c->stream));
    return CAPAZ_OK;
}

// Forward steps from the context's state with the observation gathers;
// every `every` steps (if > 0) the state is copied to ckpt + q*5*stride.
// seen: [nsamples][2][ncells] on the device.  With rec_seg != NULL the steps
// from rec_from on are run by the recording kernel into the segment buffers
// (states at rec_seg + (t - rec_from)*5*stride for t > rec_from, the
// divergence of step t at rec_div + (t - rec_from)*3*stride), the way the
// backward sweep recomputes a segment, and the context then takes the last
// two states as its current and previous ones.
int forward(capaz_ctx* c, long nsteps, const Obs& ob, const int64_t* d_cells, double* seen, long every, double* ckpt,
            size_t stride, long rec_from = -1, double* rec_seg = nullptr, double* rec_div = nullptr) {
    int rc;
    if (every > 0 && (rc = copy_state(c, ckpt, stride))) return rc;
    const long n_plain = rec_seg ? rec_from : nsteps;
    auto gather = [&](long t1, const double* Sw, const double* Sn) -> int {
        const int s = ob.sample_of[(size_t)t1];
        if (s >= 0)
            ADJ_HIP(c, capaz::launch_obs_gather(Sw, Sn, d_cells, ob.ncells, seen + (size_t)s * 2 * ob.ncells,
                                                seen + (size_t)s * 2 * ob.ncells + ob.ncells, c->stream));
        return CAPAZ_OK;
    };
    for (long t = 0; t < n_plain; ++t) {
        if ((rc = capaz_step(c, t + 1, 1))) return rc;
        const long t1 = t + 1;
        if (every > 0 && t1 % every == 0 && t1 < nsteps &&
            (rc = copy_state(c, ckpt + (size_t)(t1 / every) * kState * stride, stride)))
            return rc;
        // two phases: S_w = 1 - S_n (the S_w array is an intermediate)
        if ((rc = gather(t1, c->p.model == 3 ? c->cur[capaz::F_S_w] : nullptr, c->cur[capaz::F_S_n]))) return rc;
    }
    if (!rec_seg) return CAPAZ_OK;
    // the recorded tail starts from the checkpoint just taken (or the initial copy)
    const double* x0 = ckpt + (size_t)(rec_from / (every > 0 ? every : 1)) * kState * stride;
    auto slot = [&](long t) -> double* {
        return t == rec_from ? const_cast<double*>(x0) : rec_seg + (size_t)(t - rec_from) * kState * stride;
    };
    for (long t = rec_from; t < nsteps; ++t) {
        if ((rc = fused_between(c, c->stream, slot(t), slot(t + 1), stride,
                                rec_div + (size_t)(t - rec_from) * 3 * stride)))
            return rc;
        if ((rc = gather(t + 1, c->p.model == 3 ? slot(t + 1) + stride : nullptr, slot(t + 1) + 2 * stride)))
            return rc;
    }
    const size_t bytes = (size_t)c->p.n * sizeof(double);
    for (int f = 0; f < kState; ++f) {
        if (!state_field(c, f)) continue;
        ADJ_HIP(c, hipMemcpyAsync(c->cur[kStateField[f]], slot(nsteps) + f * stride, bytes, hipMemcpyDeviceToDevice,
                                  c->stream));
        ADJ_HIP(c, hipMemcpyAsync(c->alt[kStateField[f]], slot(nsteps - 1) + f * stride, bytes,
                                  hipMemcpyDeviceToDevice, c->stream));
    }
    c->stale = true;  // intermediates follow from the previous state, as after capaz_step
    return CAPAZ_OK;
}

// 0.5 * sum of squared mismatches, in sample-then-cell order.
int objective_value(capaz_ctx* c, const capaz_observations* o, const Obs& ob, const double* seen, double* J) {
    std::vector<double> h((size_t)ob.nsamples * 2 * ob.ncells);
    if (!h.empty())
        ADJ_HIP(c, hipMemcpyAsync(h.data(), seen, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ADJ_HIP(c, hipStreamSynchronize(c->stream));
    double sum = 0.;
    for (int s = 0; s < ob.nsamples; ++s)
        for (int64_t q = 0; q < ob.ncells; ++q) {
            const double dw = h[(size_t)s * 2 * ob.ncells + q] - o->S_w[(size_t)s * ob.ncells + q];
            const double dn = h[(size_t)s * 2 * ob.ncells + ob.ncells + q] - o->S_n[(size_t)s * ob.ncells + q];
            sum += dw * dw + dn * dn;
        }
    *J = 0.5 * sum;
    return CAPAZ_OK;
}

size_t field_stride(int64_t n) { return ((size_t)n + 31) & ~(size_t)31; }  // 256-byte aligned fields

}  // namespace

extern "C" {

int capaz_objective(capaz_ctx* c, long nsteps, const capaz_observations* o, double* J) {
    int rc;
    if (!c || !J) return CAPAZ_ERR_INVALID;
    if ((rc = capaz::detail::check_ctx(c))) return rc;
    if ((rc = check_model(c))) return rc;
    Obs ob;
    if ((rc = check_obs(c, nsteps, o, &ob))) return rc;
    if ((rc = capaz::detail::scan_m(c))) return rc;
    if (c->m_has_zero) return fail(c, CAPAZ_ERR_STATE, "adjoint: cells with m == 0");
    const size_t ns = (size_t)ob.nsamples * 2 * ob.ncells;
    double* w;
    if ((rc = workspace(c, ns + (size_t)ob.ncells, &w))) return rc;
    int64_t* d_cells = reinterpret_cast<int64_t*>(w + ns);
    if (ob.ncells)
        ADJ_HIP(c, hipMemcpyAsync(d_cells, o->cells, ob.ncells * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    if ((rc = forward(c, nsteps, ob, d_cells, w, 0, nullptr, 0))) return rc;
    return objective_value(c, o, ob, w, J);
}

int capaz_tangent(capaz_ctx* c, long nsteps, const capaz_observations* o, const double* dlnK, const double* dm,
                  double* J, double* dJ) {
    using namespace capaz;
    int rc;
    if (!c || !J || !dJ) return CAPAZ_ERR_INVALID;
    if ((rc = detail::check_ctx(c))) return rc;
    if ((rc = check_model(c))) return rc;
    if (c->p.model != 3) return fail(c, CAPAZ_ERR_INVALID, "tangent: the three-phase model only");
    Obs ob;
    if ((rc = check_obs(c, nsteps, o, &ob))) return rc;
    if ((rc = detail::scan_m(c))) return rc;
    if (c->m_has_zero) return fail(c, CAPAZ_ERR_STATE, "adjoint: cells with m == 0");
    const size_t S = field_stride(c->p.n), n = (size_t)c->p.n;
    const size_t ns = (size_t)ob.nsamples * 2 * ob.ncells;
    // layout: dx x2 (5 fields each) | dlnK, dm | seen | dseen | cells
    double* w;
    if ((rc = workspace(c, 12 * S + 2 * ns + (size_t)ob.ncells, &w))) return rc;
    double* dx[2] = {w, w + 5 * S};
    double* dth = w + 10 * S;
    double* seen = w + 12 * S;
    double* dseen = seen + ns;
    int64_t* d_cells = reinterpret_cast<int64_t*>(dseen + ns);
    ADJ_HIP(c, hipMemsetAsync(w, 0, 12 * S * sizeof(double), c->stream));
    if (dlnK) ADJ_HIP(c, hipMemcpyAsync(dth, dlnK, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (dm) ADJ_HIP(c, hipMemcpyAsync(dth + S, dm, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (ob.ncells)
        ADJ_HIP(c, hipMemcpyAsync(d_cells, o->cells, ob.ncells * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    int cur = 0;
    for (long t = 0; t < nsteps; ++t) {
        // the tangent of step t reads x_t: queued before the step itself
        TanIO a;
        a.Pw = c->cur[F_P_w];
        a.Sw = c->cur[F_S_w];
        a.Sn = c->cur[F_S_n];
        a.Pn = c->cur[F_P_n];
        a.Pg = c->cur[F_P_g];
        a.m = c->cur[F_m];
        a.K = c->cur[F_K];
        a.dtm = c->d_dtm;
        const double* d0 = dx[cur];
        double* d1 = dx[cur ^ 1];
        a.dPw = d0;
        a.dSw = d0 + S;
        a.dSn = d0 + 2 * S;
        a.dPn = d0 + 3 * S;
        a.dPg = d0 + 4 * S;
        a.dlnK = dth;
        a.dm = dth + S;
        a.dPw1 = d1;
        a.dSw1 = d1 + S;
        a.dSn1 = d1 + 2 * S;
        a.dPn1 = d1 + 3 * S;
        a.dPg1 = d1 + 4 * S;
        ADJ_HIP(c, launch_tangent_step(c->p, a, c->stream));
        if ((rc = capaz_step(c, t + 1, 1))) return rc;
        cur ^= 1;
        const int smp = ob.sample_of[(size_t)(t + 1)];
        if (smp >= 0) {
            const size_t off = (size_t)smp * 2 * ob.ncells;
            ADJ_HIP(c, launch_obs_gather(c->cur[F_S_w], c->cur[F_S_n], d_cells, ob.ncells, seen + off,
                                         seen + off + ob.ncells, c->stream));
            ADJ_HIP(c, launch_obs_gather(dx[cur] + S, dx[cur] + 2 * S, d_cells, ob.ncells, dseen + off,
                                         dseen + off + ob.ncells, c->stream));
        }
    }
    if ((rc = objective_value(c, o, ob, seen, J))) return rc;
    std::vector<double> hs(ns), hd(ns);
    if (ns) {
        ADJ_HIP(c, hipMemcpy(hs.data(), seen, ns * sizeof(double), hipMemcpyDeviceToHost));
        ADJ_HIP(c, hipMemcpy(hd.data(), dseen, ns * sizeof(double), hipMemcpyDeviceToHost));
    }
    double sum = 0.;
    for (int s = 0; s < ob.nsamples; ++s)
        for (int64_t q = 0; q < ob.ncells; ++q) {
            const size_t b = (size_t)s * 2 * ob.ncells;
            sum += (hs[b + q] - o->S_w[(size_t)s * ob.ncells + q]) * hd[b + q] +
                   (hs[b + ob.ncells + q] - o->S_n[(size_t)s * ob.ncells + q]) * hd[b + ob.ncells + q];
        }
    *dJ = sum;
    return CAPAZ_OK;
}

int capaz_adjoint(capaz_ctx* c, long nsteps, const capaz_observations* o, int checkpoint_every, double* J,
                  double* grad_lnK, double* grad_m) {
    using namespace capaz;
    int rc;
    if (!c || !J) return CAPAZ_ERR_INVALID;
    if ((rc = detail::check_ctx(c))) return rc;
    if ((rc = check_model(c))) return rc;
    Obs ob;
    if ((rc = check_obs(c, nsteps, o, &ob))) return rc;
    if ((rc = detail::scan_m(c))) return rc;
    if (c->m_has_zero) return fail(c, CAPAZ_ERR_STATE, "adjoint: cells with m == 0");
    const size_t S = field_stride(c->p.n);
    const size_t n_state = (size_t)kState * S;
    // layout: checkpoints | segment buffers (1, or 2 with two streams and
    // more than one segment): states (C + 1) | recorded divergence (C x 3) |
    // lambda x2 | work (6) | grads (2) | seen | observed | cells
    const bool recorded = c->tuning.adjoint_kernel == 2;
    const size_t ns = (size_t)ob.nsamples * 2 * ob.ncells;
    const size_t n_work = 6 * S;
    auto buffers_for = [&](long seg) { return c->tuning.adjoint_streams == 2 && seg < nsteps ? 2 : 1; };
    auto doubles_for = [&](long seg) {
        const long segs = (nsteps + seg - 1) / seg;
        const size_t per_buf = (size_t)(seg + 1) * n_state + (recorded ? (size_t)seg * 3 * S : 0);
        return (size_t)segs * n_state + (size_t)buffers_for(seg) * per_buf + 2 * n_state + n_work + 2 * S + 2 * ns +
               (size_t)ob.ncells;
    };
    // Segment length C: checkpoint_every if > 0; round(sqrt(nsteps)) if < 0;
    // if 0, the whole trajectory as one segment (recorded by the forward
    // sweep, nothing recomputed) when that takes at most a quarter of the
    // device's memory and fits in what is free, else round(sqrt(nsteps)).
    // Every C gives the same bits.
    const long sqrt_c = std::max(1L, std::min((long)std::lround(std::sqrt((double)nsteps)), nsteps));
    long C = checkpoint_every > 0 ? std::max(1L, std::min((long)checkpoint_every, nsteps)) : sqrt_c;
    bool whole = false;
    if (checkpoint_every == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const size_t need = doubles_for(nsteps) * sizeof(double);
            whole = need <= total_b / 4 && need + ((size_t)1 << 30) <= free_b + c->adj_bytes;
        } else {
            (void)hipGetLastError();
        }
        if (whole) C = nsteps;
    }
    double* w;
    rc = workspace(c, doubles_for(C), &w);
    if (rc == CAPAZ_ERR_NOMEM && whole) {
        // fall back to sqrt(N) checkpoints
        (void)hipGetLastError();
        C = sqrt_c;
        rc = workspace(c, doubles_for(C), &w);
    }
    if (rc) return rc;
    c->adj_segment = C;
    const long nseg = (nsteps + C - 1) / C;
    const int nbuf = buffers_for(C);
    const size_t n_div = recorded ? (size_t)C * 3 * S : 0;
    double* ckpt = w;
    double* seg[2];  // segment buffer b: C + 1 states, then C recorded divergences
    double* divs[2];
    seg[0] = ckpt + (size_t)nseg * n_state;
    divs[0] = seg[0] + (size_t)(C + 1) * n_state;
    seg[1] = divs[0] + n_div;
    divs[1] = seg[1] + (size_t)(C + 1) * n_state;
    double* const bufs_end = nbuf == 2 ? divs[1] + n_div : seg[1];
    double* lam[2] = {bufs_end, bufs_end + n_state};
    double* work = lam[1] + n_state;
    double* grad = work + n_work;
    double* seen = grad + 2 * S;
    double* d_want = seen + ns;
    int64_t* d_cells = reinterpret_cast<int64_t*>(d_want + ns);
    if (ob.ncells)
        ADJ_HIP(c, hipMemcpyAsync(d_cells, o->cells, ob.ncells * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    std::vector<double> want(ns);  // the observed values, laid out as `seen`
    for (int s = 0; s < ob.nsamples; ++s)
        for (int64_t q = 0; q < ob.ncells; ++q) {
            want[(size_t)s * 2 * ob.ncells + q] = o->S_w[(size_t)s * ob.ncells + q];
            want[(size_t)s * 2 * ob.ncells + ob.ncells + q] = o->S_n[(size_t)s * ob.ncells + q];
        }
    if (ns) ADJ_HIP(c, hipMemcpyAsync(d_want, want.data(), ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // forward sweep (objective_value waits for it, so `want` outlives its copy);
    // recording, it runs the last segment the way the backward sweep would
    const long last0 = (nseg - 1) * C;
    const int last_buf = (int)((nseg - 1) % nbuf);
    if ((rc = recorded ? forward(c, nsteps, ob, d_cells, seen, C, ckpt, S, last0, seg[last_buf], divs[last_buf])
                       : forward(c, nsteps, ob, d_cells, seen, C, ckpt, S)))
        return rc;
    if ((rc = objective_value(c, o, ob, seen, J))) return rc;
    auto obs_adjoint = [&](double* l, long step) -> int {
        const int s = ob.sample_of[(size_t)step];
        if (s < 0) return CAPAZ_OK;
        const size_t off = (size_t)s * 2 * ob.ncells;
        hipError_t e = launch_obs_adjoint(c->p.model == 3 ? l + 1 * S : nullptr, l + 2 * S, d_cells, ob.ncells,
                                          seen + off, seen + off + ob.ncells, d_want + off, d_want + off + ob.ncells,
                                          c->stream);
        return e == hipSuccess ? CAPAZ_OK : hip_fail(c, e, "launch_obs_adjoint");
    };
    // reverse sweep
    hipError_t e;
    if ((e = hipMemsetAsync(lam[0], 0, n_state * sizeof(double), c->stream)) != hipSuccess ||
        (e = hipMemsetAsync(grad, 0, 2 * S * sizeof(double), c->stream)) != hipSuccess)
        return hip_fail(c, e, "hipMemsetAsync");
    if ((rc = obs_adjoint(lam[0], nsteps))) return rc;
    auto io_at = [&](const double* x, double* w, const double* l1, double* l0) {
        AdjIO a;
        a.Pw = x;
        a.Sw = x + S;
        a.Sn = x + 2 * S;
        a.Pn = x + 3 * S;
        a.Pg = x + 4 * S;
        a.m = c->cur[F_m];
        a.K = c->cur[F_K];
        a.dtm = c->d_dtm;
        a.lPw1 = l1;
        a.lSw1 = l1 + S;
        a.lSn1 = l1 + 2 * S;
        a.lPn1 = l1 + 3 * S;
        a.lPg1 = l1 + 4 * S;
        a.mu_w = w;
        a.mu_n = w + S;
        a.mu_g = w + 2 * S;
        a.dPw = w + 3 * S;
        a.dSw = w + 4 * S;
        a.dSn = w + 5 * S;
        a.lPw = l0;
        a.lSw = l0 + S;
        a.lSn = l0 + 2 * S;
        a.lPn = l0 + 3 * S;
        a.lPg = l0 + 4 * S;
        a.g_lnK = grad;
        a.g_m = grad + S;
        return a;
    };
    // Segment q lives in buffer q % nbuf.  With two buffers the recompute of
    // segment q - 1 runs on the second stream while the passes reverse
    // segment q; events order each recompute after the passes that last read
    // its buffer, and the passes after the recompute of their segment.  The
    // kernels and their inputs are those of one stream: the bits are the same.
    hipStream_t s2 = c->stream;
    if (nbuf == 2) {
        if (!c->adj_stream) ADJ_HIP(c, hipStreamCreateWithFlags(&c->adj_stream, hipStreamNonBlocking));
        for (hipEvent_t& ev : c->adj_ev)
            if (!ev) ADJ_HIP(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        s2 = c->adj_stream;
        // both buffers free once the forward sweep and the memsets are done
        for (int b = 0; b < 2; ++b) ADJ_HIP(c, hipEventRecord(c->adj_ev[2 + b], c->stream));
    }
    // the segment's states (and, recording, every step's divergence: the
    // last step's new state lands in the spare slot C)
    auto slot_of = [&](long q, long t) {
        const long t0 = q * C;
        return t == t0 ? ckpt + (size_t)q * n_state : seg[q % nbuf] + (size_t)(t - t0) * n_state;
    };
    auto recompute = [&](long q, hipStream_t s) -> int {
        const long t0 = q * C, t1 = std::min(nsteps, t0 + C);
        double* dv = divs[q % nbuf];
        for (long t = t0; t + 1 < t1 || (recorded && t < t1); ++t)
            if ((rc = fused_between(c, s, slot_of(q, t), slot_of(q, t + 1), S,
                                    recorded ? dv + (size_t)(t - t0) * 3 * S : nullptr)))
                return rc;
        return CAPAZ_OK;
    };
    if (!recorded && (rc = recompute(nseg - 1, c->stream))) return rc;  // else recorded by the forward sweep
    int cur = 0;
    for (long q = nseg - 1; q >= 0; --q) {
        const long t0 = q * C, t1 = std::min(nsteps, t0 + C);
        if (q >= 1 && nbuf == 2) {
            const int b = (int)((q - 1) % 2);
            ADJ_HIP(c, hipStreamWaitEvent(s2, c->adj_ev[2 + b], 0));
            if ((rc = recompute(q - 1, s2))) return rc;
            ADJ_HIP(c, hipEventRecord(c->adj_ev[b], s2));
        }
        double* dv = divs[q % nbuf];
        for (long t = t1 - 1; t >= t0; --t) {
            double* l0 = lam[cur ^ 1];
            const AdjIO a = io_at(slot_of(q, t), work, lam[cur], l0);
            if (recorded)
                e = launch_adjoint_step_recorded(c->p, a, dv + (size_t)(t - t0) * 3 * S, (int64_t)S, c->stream);
            else
                e = c->tuning.adjoint_kernel == 1 ? launch_adjoint_step_tiled(c->p, a, c->stream)
                                                  : launch_adjoint_step(c->p, a, c->stream);
            if (e != hipSuccess)
                return hip_fail(c, e, "launch_adjoint_step");
            if (t >= 1 && (rc = obs_adjoint(l0, t))) return rc;
            cur ^= 1;
        }
        if (q >= 1 && nbuf == 2) {
            ADJ_HIP(c, hipEventRecord(c->adj_ev[2 + q % 2], c->stream));
            ADJ_HIP(c, hipStreamWaitEvent(c->stream, c->adj_ev[(q - 1) % 2], 0));
        } else if (q >= 1 && (rc = recompute(q - 1, c->stream))) {
            return rc;
        }
    }
    const size_t bytes = (size_t)c->p.n * sizeof(double);
    if (grad_lnK && (e = hipMemcpyAsync(grad_lnK, grad, bytes, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
        return hip_fail(c, e, "hipMemcpyAsync");
    if (grad_m && (e = hipMemcpyAsync(grad_m, grad + S, bytes, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
        return hip_fail(c, e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return hip_fail(c, e, "hipStreamSynchronize");
    return CAPAZ_OK;
}

}  // extern "C"
