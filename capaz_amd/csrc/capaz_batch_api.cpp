// C ABI of batched ensembles (include/capaz.h, capaz_batch_*): many
// realizations of one physics, each a full single-domain context (its own
// grid, fields, entry points and fused step), advanced together by the batch
// kernels of capaz_batch.hip.  A run of n steps is one HIP graph of n steps
// (interior tiles + boundary cells per step), captured once per step count
// and tile shape and replayed; the table of realizations it reads is
// refreshed before every launch, so the graph stays valid while members
// swap their ping-pong arrays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "capaz_batch.h"
#include "capaz_ctx.h"

struct capaz_batch {
    capaz_params base;
    std::vector<capaz_ctx*> dom;
    hipStream_t stream = nullptr;
    hipEvent_t sync_ev = nullptr;
    int cu_count = 0;
    int closure_id = 0;  // 0 reference, 1 Corey with a run-time exponent, 2 Corey-2
    capaz::BatchShape shape;
    int use_graph = 1;
    // launch tables (rebuilt when the shape changes)
    std::vector<int> order;  // realizations, largest interior first
    capaz::BatchDomain* h_doms = nullptr;  // pinned
    capaz::BatchDomain* d_doms = nullptr;
    int* d_tile_owner = nullptr;
    int* d_face_owner = nullptr;
    int ntiles = 0, nface_blocks = 0, grid = 0;
    bool tables_for_shape = false;
    capaz::BatchShape table_shape;
    // persistent queue counters, one per step of a graph
    int* d_counters = nullptr;
    int ncounters = 0;
    // captured graph
    hipGraphExec_t graph = nullptr;
    long graph_steps = -1;
    std::string err;
};

namespace {

thread_local std::string g_batch_create_error;

int bfail(capaz_batch* b, int code, const std::string& msg) {
    if (b)
        b->err = msg;
    else
        g_batch_create_error = msg;
    return code;
}

int bhip(capaz_batch* b, hipError_t e, const char* what) {
    return bfail(b, e == hipErrorOutOfMemory ? CAPAZ_ERR_NOMEM : CAPAZ_ERR_DEVICE,
                 std::string(what) + ": " + hipGetErrorString(e));
}

#define BATCH_HIP(b, call)                                 \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return bhip((b), e_, #call); \
    } while (0)

int64_t interior_cells(const capaz::DevParams& p) {
    return (int64_t)(p.Nx >= 3 ? p.Nx - 2 : 1) * (p.Ny - 2) * (p.Nz >= 3 ? p.Nz - 2 : 1);
}

void drop_graph(capaz_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_steps = -1;
}

// Tile and face tables of the current shape (device copies, sizes).
int build_tables(capaz_batch* b) {
    const auto& sh = b->shape;
    if (!capaz::batch_shape_supported(sh.tx, sh.ty))
        return bfail(b, CAPAZ_ERR_INVALID, "capaz_batch: tile shapes are 64x8, 64x4 and 32x8");
    if (b->tables_for_shape && b->table_shape.tx == sh.tx && b->table_shape.ty == sh.ty &&
        b->table_shape.kchunk == sh.kchunk && b->table_shape.persistent == sh.persistent)
        return CAPAZ_OK;
    drop_graph(b);
    BATCH_HIP(b, hipStreamSynchronize(b->stream));
    const int n = (int)b->dom.size();
    std::vector<int> tile_owner, face_owner;
    int64_t tiles = 0, fblocks = 0;
    for (int slot = 0; slot < n; ++slot) {
        capaz::BatchDomain& d = b->h_doms[slot];
        const capaz::DevParams& p = b->dom[b->order[slot]]->p;
        capaz::batch_tiles(p, sh, &d.tiles_x, &d.tiles_y, &d.tiles_z);
        const int64_t t = (int64_t)d.tiles_x * d.tiles_y * d.tiles_z;
        const int64_t fc = capaz::batch_face_cells(p);
        d.tile_begin = (int)tiles;
        d.face_block_begin = (int)fblocks;
        d.face_blocks = (int)((fc + 255) / 256);
        tiles += t;
        fblocks += d.face_blocks;
        if (tiles > (1LL << 30) || fblocks > (1LL << 30))
            return bfail(b, CAPAZ_ERR_INVALID, "capaz_batch: too many tiles for one launch");
        tile_owner.insert(tile_owner.end(), (size_t)t, slot);
        face_owner.insert(face_owner.end(), (size_t)d.face_blocks, slot);
    }
    if (b->d_tile_owner) (void)hipFree(b->d_tile_owner);
    if (b->d_face_owner) (void)hipFree(b->d_face_owner);
    b->d_tile_owner = b->d_face_owner = nullptr;
    BATCH_HIP(b, hipMalloc(&b->d_tile_owner, std::max<size_t>(1, tile_owner.size()) * sizeof(int)));
    BATCH_HIP(b, hipMalloc(&b->d_face_owner, std::max<size_t>(1, face_owner.size()) * sizeof(int)));
    BATCH_HIP(b, hipMemcpy(b->d_tile_owner, tile_owner.data(), tile_owner.size() * sizeof(int), hipMemcpyHostToDevice));
    BATCH_HIP(b, hipMemcpy(b->d_face_owner, face_owner.data(), face_owner.size() * sizeof(int), hipMemcpyHostToDevice));
    b->ntiles = (int)tiles;
    b->nface_blocks = (int)fblocks;
    int per_cu = 0;
    BATCH_HIP(b, capaz::batch_occupancy(b->base.model, b->closure_id, sh, &per_cu));
    b->grid = std::max(1, std::min(b->ntiles, std::max(1, per_cu) * b->cu_count));
    b->table_shape = sh;
    b->tables_for_shape = true;
    return CAPAZ_OK;
}

// Member contexts ready for a batch launch, and their current arrays in the
// table (uploaded on the batch stream after the members' queued work).
int prepare(capaz_batch* b) {
    using namespace capaz;
    int rc;
    if ((rc = build_tables(b))) return rc;
    BATCH_HIP(b, hipStreamSynchronize(b->stream));  // the pinned table is free to rewrite
    const int n = (int)b->dom.size();
    for (int slot = 0; slot < n; ++slot) {
        capaz_ctx* c = b->dom[b->order[slot]];
        if ((rc = detail::scan_m(c))) return bfail(b, rc, std::string("member: ") + c->err);
        if (c->m_has_zero)
            return bfail(b, CAPAZ_ERR_STATE, "capaz_batch: a realization has cells with m == 0 (use its context's "
                                             "capaz_step_sequence)");
        BatchDomain& d = b->h_doms[slot];
        d.p = c->p;
        StepIO fw;
        fw.Pw = c->cur[F_P_w];
        fw.Sw = c->cur[F_S_w];
        fw.Sn = c->cur[F_S_n];
        fw.Pn = c->cur[F_P_n];
        fw.Pg = c->cur[F_P_g];
        fw.m = c->cur[F_m];
        fw.K = c->cur[F_K];
        fw.dtm = c->d_dtm;
        fw.Pw1 = c->alt[F_P_w];
        fw.Sw1 = c->alt[F_S_w];
        fw.Sn1 = c->alt[F_S_n];
        fw.Pn1 = c->alt[F_P_n];
        fw.Pg1 = c->alt[F_P_g];
        StepIO bw = fw;  // the second arrays back into the current ones
        bw.Pw = c->alt[F_P_w];
        bw.Sw = c->alt[F_S_w];
        bw.Sn = c->alt[F_S_n];
        bw.Pn = c->alt[F_P_n];
        bw.Pg = c->alt[F_P_g];
        bw.Pw1 = c->cur[F_P_w];
        bw.Sw1 = c->cur[F_S_w];
        bw.Sn1 = c->cur[F_S_n];
        bw.Pn1 = c->cur[F_P_n];
        bw.Pg1 = c->cur[F_P_g];
        d.io[0] = fw;
        d.io[1] = bw;
        // the batch stream waits for the member's queued work
        BATCH_HIP(b, hipEventRecord(b->sync_ev, c->stream));
        BATCH_HIP(b, hipStreamWaitEvent(b->stream, b->sync_ev, 0));
    }
    BATCH_HIP(b, hipMemcpyAsync(b->d_doms, b->h_doms, n * sizeof(BatchDomain), hipMemcpyHostToDevice, b->stream));
    return CAPAZ_OK;
}

int ensure_counters(capaz_batch* b, long nsteps) {
    if (b->ncounters >= nsteps) return CAPAZ_OK;
    drop_graph(b);
    BATCH_HIP(b, hipStreamSynchronize(b->stream));
    if (b->d_counters) (void)hipFree(b->d_counters);
    b->d_counters = nullptr;
    b->ncounters = 0;
    BATCH_HIP(b, hipMalloc(&b->d_counters, (size_t)nsteps * sizeof(int)));
    b->ncounters = (int)nsteps;
    return CAPAZ_OK;
}

// nsteps batch steps as plain launches on the batch stream.
int enqueue_steps(capaz_batch* b, long nsteps) {
    if (b->shape.persistent) BATCH_HIP(b, capaz::launch_zero_ints(b->d_counters, (int)nsteps, b->stream));
    for (long s = 0; s < nsteps; ++s)
        BATCH_HIP(b, capaz::launch_batch_step(b->base.model, b->closure_id, b->shape, b->d_doms, b->d_tile_owner,
                                              b->ntiles, b->d_face_owner, b->nface_blocks, (int)(s & 1),
                                              b->d_counters + s, b->grid, b->stream));
    return CAPAZ_OK;
}

// After nsteps batch steps: members swap their arrays (odd counts), mark
// their intermediates stale and order their streams after the batch.
int finish(capaz_batch* b, long nsteps) {
    BATCH_HIP(b, hipEventRecord(b->sync_ev, b->stream));
    for (capaz_ctx* c : b->dom) {
        if (nsteps & 1)
            for (int f = 0; f < capaz::F_COUNT; ++f)
                if (c->primary[f]) std::swap(c->cur[f], c->alt[f]);
        if (nsteps > 0) c->stale = true;
        BATCH_HIP(b, hipStreamWaitEvent(c->stream, b->sync_ev, 0));
    }
    return CAPAZ_OK;
}

int run_steps(capaz_batch* b, long nsteps, bool graph) {
    int rc;
    if (nsteps == 0) return CAPAZ_OK;
    if ((rc = ensure_counters(b, nsteps))) return rc;
    if ((rc = prepare(b))) return rc;
    if (graph) {
        if (!b->graph || b->graph_steps != nsteps) {
            drop_graph(b);
            hipGraph_t g = nullptr;
            BATCH_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_steps(b, nsteps);
            hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return bhip(b, e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->graph = nullptr;
                return bhip(b, e, "hipGraphInstantiate");
            }
            b->graph_steps = nsteps;
        }
        BATCH_HIP(b, hipGraphLaunch(b->graph, b->stream));
    } else if ((rc = enqueue_steps(b, nsteps))) {
        return rc;
    }
    return finish(b, nsteps);
}

int check_batch(capaz_batch* b) {
    if (!b) return CAPAZ_ERR_INVALID;
    BATCH_HIP(b, hipSetDevice(b->base.device));
    return CAPAZ_OK;
}

}  // namespace

extern "C" {

int capaz_batch_create(const capaz_params* base, int count, const int* dims, capaz_batch** out) {
    if (!base || !dims || !out || count < 1) return bfail(nullptr, CAPAZ_ERR_INVALID, "capaz_batch_create: bad argument");
    *out = nullptr;
    if (base->energy) return bfail(nullptr, CAPAZ_ERR_INVALID, "capaz_batch_create: no batched energy step");
    capaz_batch* b = new (std::nothrow) capaz_batch;
    if (!b) return bfail(nullptr, CAPAZ_ERR_NOMEM, "host allocation failed");
    b->base = *base;
    auto bail = [&](int code, const std::string& msg) {
        capaz_batch_destroy(b);
        return bfail(nullptr, code, msg);
    };
    for (int r = 0; r < count; ++r) {
        capaz_params p = *base;
        p.Nx = dims[3 * r];
        p.Ny = dims[3 * r + 1];
        p.Nz = dims[3 * r + 2];
        capaz_ctx* c = nullptr;
        int rc = capaz_create(&p, &c);
        if (rc) return bail(rc, "realization " + std::to_string(r) + ": " + capaz_last_error(nullptr));
        b->dom.push_back(c);
    }
    const capaz::DevParams& p0 = b->dom[0]->p;
    b->closure_id = p0.closure == 0 ? 0 : (p0.corey_n == 2. ? 2 : 1);
    hipError_t e;
    if ((e = hipSetDevice(base->device)) != hipSuccess) return bail(CAPAZ_ERR_DEVICE, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(CAPAZ_ERR_DEVICE, "hipStreamCreate");
    if ((e = hipEventCreateWithFlags(&b->sync_ev, hipEventDisableTiming)) != hipSuccess)
        return bail(CAPAZ_ERR_DEVICE, "hipEventCreate");
    if ((e = hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, base->device)) != hipSuccess)
        return bail(CAPAZ_ERR_DEVICE, "hipDeviceGetAttribute");
    if ((e = hipHostMalloc(&b->h_doms, count * sizeof(capaz::BatchDomain))) != hipSuccess)
        return bail(CAPAZ_ERR_NOMEM, "hipHostMalloc");
    std::memset(b->h_doms, 0, count * sizeof(capaz::BatchDomain));
    if ((e = hipMalloc(&b->d_doms, count * sizeof(capaz::BatchDomain))) != hipSuccess)
        return bail(CAPAZ_ERR_NOMEM, "hipMalloc");
    b->order.resize(count);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        return interior_cells(b->dom[x]->p) > interior_cells(b->dom[y]->p);
    });
    *out = b;
    return CAPAZ_OK;
}

int capaz_batch_destroy(capaz_batch* b) {
    if (!b) return CAPAZ_OK;
    (void)hipSetDevice(b->base.device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_tile_owner) (void)hipFree(b->d_tile_owner);
    if (b->d_face_owner) (void)hipFree(b->d_face_owner);
    if (b->d_doms) (void)hipFree(b->d_doms);
    if (b->h_doms) (void)hipHostFree(b->h_doms);
    if (b->sync_ev) (void)hipEventDestroy(b->sync_ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (capaz_ctx* c : b->dom) capaz_destroy(c);
    delete b;
    return CAPAZ_OK;
}

const char* capaz_batch_last_error(const capaz_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int capaz_batch_size(const capaz_batch* b) { return b ? (int)b->dom.size() : 0; }

int capaz_batch_domain(capaz_batch* b, int r, capaz_ctx** ctx) {
    if (!b || !ctx) return CAPAZ_ERR_INVALID;
    if (r < 0 || r >= (int)b->dom.size()) return bfail(b, CAPAZ_ERR_INVALID, "capaz_batch_domain: index out of range");
    *ctx = b->dom[r];
    return CAPAZ_OK;
}

int capaz_batch_step(capaz_batch* b, long first_step, long nsteps) {
    (void)first_step;  // the fused step does not read t (assign_roS_nr, cpu.cpp:619-731)
    int rc;
    if ((rc = check_batch(b))) return rc;
    if (nsteps < 0 || nsteps > (1L << 24)) return bfail(b, CAPAZ_ERR_INVALID, "capaz_batch_step: nsteps out of range");
    for (capaz_ctx* c : b->dom)
        if (c->halo_pack)
            return bfail(b, CAPAZ_ERR_INVALID, "capaz_batch_step: a member has a halo pack, which batched steps do not write");
    return run_steps(b, nsteps, b->use_graph != 0);
}

int capaz_batch_tune(capaz_batch* b, const char* key, int value) {
    if (!b || !key) return CAPAZ_ERR_INVALID;
    const std::string k(key);
    capaz::BatchShape sh = b->shape;
    if (k == "tile_x") sh.tx = value;
    else if (k == "tile_y") sh.ty = value;
    else if (k == "kchunk" && value >= 1 && value <= 1 << 20) sh.kchunk = value;
    else if (k == "persistent" && (value == 0 || value == 1)) sh.persistent = value;
    else if (k == "graph" && (value == 0 || value == 1)) {
        b->use_graph = value;
        return CAPAZ_OK;
    } else {
        return bfail(b, CAPAZ_ERR_INVALID, "capaz_batch_tune: unknown key or value out of range");
    }
    if (k == "tile_x" || k == "tile_y") {
        // the pair is checked when both are set: accept an intermediate state
        // only if some supported shape has this value
        const bool ok = k == "tile_x" ? (value == 64 || value == 32) : (value == 8 || value == 4);
        if (!ok) return bfail(b, CAPAZ_ERR_INVALID, "capaz_batch_tune: tile shapes are 64x8, 64x4 and 32x8");
    }
    b->shape = sh;
    return CAPAZ_OK;
}

int capaz_batch_get_tuning(const capaz_batch* b, const char* key, int* value) {
    if (!b || !key || !value) return CAPAZ_ERR_INVALID;
    const std::string k(key);
    if (k == "tile_x") *value = b->shape.tx;
    else if (k == "tile_y") *value = b->shape.ty;
    else if (k == "kchunk") *value = b->shape.kchunk;
    else if (k == "persistent") *value = b->shape.persistent;
    else if (k == "graph") *value = b->use_graph;
    else return CAPAZ_ERR_INVALID;
    return CAPAZ_OK;
}

int capaz_batch_autotune(capaz_batch* b, int steps, double* best_ms) {
    using namespace capaz;
    int rc;
    if ((rc = check_batch(b))) return rc;
    if (steps < 2) steps = 2;
    steps += steps & 1;  // even: the members' current arrays stay in place
    // Intermediates current, so the second arrays are free to overwrite; keep
    // a copy of the primary state and restore it after every trial.
    std::vector<double*> saved;
    auto release = [&]() {
        for (double* d : saved) (void)hipFree(d);
    };
    for (capaz_ctx* c : b->dom) {
        if ((rc = detail::materialize(c))) {
            release();
            return bfail(b, rc, std::string("member: ") + c->err);
        }
        for (int f = 0; f < F_COUNT; ++f) {
            if (!c->primary[f]) continue;
            double* d = nullptr;
            hipError_t e = hipMalloc(&d, (size_t)c->p.n * sizeof(double));
            if (e != hipSuccess) {
                release();
                return bhip(b, e, "hipMalloc");
            }
            saved.push_back(d);
            e = hipMemcpyAsync(d, c->cur[f], (size_t)c->p.n * sizeof(double), hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) {
                release();
                return bhip(b, e, "hipMemcpyAsync");
            }
        }
    }
    auto restore = [&]() -> int {
        BATCH_HIP(b, hipStreamSynchronize(b->stream));
        size_t q = 0;
        for (capaz_ctx* c : b->dom)
            for (int f = 0; f < F_COUNT; ++f)
                if (c->primary[f]) {
                    BATCH_HIP(b, hipMemcpyAsync(c->cur[f], saved[q++], (size_t)c->p.n * sizeof(double),
                                                hipMemcpyDeviceToDevice, c->stream));
                }
        for (capaz_ctx* c : b->dom) {
            BATCH_HIP(b, hipStreamSynchronize(c->stream));
            c->stale = false;
        }
        return CAPAZ_OK;
    };
    static const int shapes[][2] = {{64, 8}, {64, 4}, {32, 8}};
    static const int kchunks[] = {8, 16, 32};
    const BatchShape keep = b->shape;
    BatchShape best = keep;
    float best_t = -1.f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = CAPAZ_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = bfail(b, CAPAZ_ERR_DEVICE, "hipEventCreate");
    for (int persistent = 1; persistent >= 0 && !rc; --persistent)
        for (const auto& s : shapes) {
            if (rc) break;
            for (int kc : kchunks) {
                b->shape = BatchShape{s[0], s[1], kc, persistent};
                if ((rc = run_steps(b, 2, false))) break;  // warm-up
                if ((rc = prepare(b))) break;
                if ((rc = ensure_counters(b, steps))) break;
                if (hipEventRecord(e0, b->stream) != hipSuccess) { rc = bfail(b, CAPAZ_ERR_DEVICE, "hipEventRecord"); break; }
                if ((rc = enqueue_steps(b, steps))) break;
                if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                    rc = bfail(b, CAPAZ_ERR_DEVICE, "hipEventSynchronize");
                    break;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if ((rc = finish(b, steps))) break;
                if ((rc = restore())) break;
                if (best_t < 0.f || ms < best_t) {
                    best_t = ms;
                    best = b->shape;
                }
            }
        }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        (void)restore();
        b->shape = keep;
    } else {
        b->shape = best;
        if (best_ms) *best_ms = (double)best_t / steps;
    }
    for (capaz_ctx* c : b->dom) (void)hipStreamSynchronize(c->stream);
    release();
    return rc;
}

int capaz_batch_sync(capaz_batch* b) {
    int rc;
    if ((rc = check_batch(b))) return rc;
    BATCH_HIP(b, hipStreamSynchronize(b->stream));
    for (capaz_ctx* c : b->dom) BATCH_HIP(b, hipStreamSynchronize(c->stream));
    return CAPAZ_OK;
}

int capaz_batch_stream(capaz_batch* b, void** s) {
    if (!b || !s) return CAPAZ_ERR_INVALID;
    *s = (void*)b->stream;
    return CAPAZ_OK;
}

}  // extern "C"
