/* capaz_amd: MI355X (gfx950) backend for capaz's explicit 3-D two- and
 * three-phase filtration step, optionally with the energy equation.
 *
 * This is the drop-in boundary: a C ABI (plain pointers and sizes, no C++ or
 * torch types) that replaces the device half of the reference's compute
 * backend, gpu.cu, behind the entry points its driver calls through
 * defines.h.  Each entry point below names the reference function it stands
 * in for.  A C++ adapter with the exact defines.h signatures
 * (capaz_amd/adapter/capaz_gpu_backend.cpp) forwards to these functions, so a
 * driver with main.cpp's structure links against this library the way it
 * linked against gpu.o.
 *
 * Storage: one FP64 array per field on the device, x fastest,
 * local = i + j*Nx + k*Nx*Ny (cpu.cpp:145), j = 0 the top face.
 *
 * Errors: every function returns CAPAZ_OK (0) or a negative capaz_status;
 * capaz_last_error() gives the message of the last failure on a context (or,
 * for a NULL context, of the last failed capaz_create on this thread).  The
 * reference's functions return void and exit(1) on error (shared_test.cpp:23-33);
 * this ABI reports instead.
 *
 * Threading: a context is used by one host thread at a time; different
 * contexts are independent.  All work of a context is ordered on its own HIP
 * stream; functions that return data to the host synchronise that stream. */
#ifndef CAPAZ_H
#define CAPAZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAPAZ_ABI_VERSION 1

typedef struct capaz_ctx capaz_ctx;

typedef enum capaz_status {
    CAPAZ_OK = 0,
    CAPAZ_ERR_INVALID = -1, /* bad argument or unsupported configuration */
    CAPAZ_ERR_DEVICE = -2,  /* a HIP call failed */
    CAPAZ_ERR_NOMEM = -3,   /* device allocation failed */
    CAPAZ_ERR_STATE = -4,   /* call not valid in the context's current state */
    CAPAZ_ERR_NONFINITE = -5 /* the state holds a NaN or an infinity (capaz_check_finite) */
} capaz_status;

typedef enum capaz_model {
    CAPAZ_TWO_PHASE = 2,  /* water / NAPL   (Two-phase/two-phase.cpp) */
    CAPAZ_THREE_PHASE = 3 /* water / NAPL / gas (Three-phase/three-phase.cpp) */
} capaz_model;

typedef enum capaz_closure {
    /* The reference's CPU closures: Brooks-Corey k_r and P_c for two phases
     * (two-phase.cpp:4-23); van Genuchten-type k_r and P_c = 0 for three
     * phases (three-phase.cpp:11-186). */
    CAPAZ_CLOSURE_REFERENCE = 0,
    /* Corey k_r = S_e^n of the effective saturations clipped to [0, 1], no
     * capillary pressure.  No reference counterpart. */
    CAPAZ_CLOSURE_COREY = 1
} capaz_closure;

typedef enum capaz_energy_closure {
    /* energy.cpp's constants and closures, with its two z-term quirks
     * (energy.cpp:257, 300-305); three phases only. */
    CAPAZ_ENERGY_REFERENCE = 0,
    /* The configurations' closure, for two or three phases: energy.cpp's
     * fluid densities, heat capacities and enthalpies (the water enthalpy
     * as the integral of its c_w), a constant conductivity lambda_eff and
     * rock heat capacity c_r, mu_a(T) = mu_a exp(-mu_T_a (T - mu_T_ref)),
     * fixed top/bottom temperatures, the conduction summed over the three
     * axes and the z advection with u_z.  The two-phase energy model's
     * Newton is energy.cpp:370-463's system without the gas row and
     * column (4 x 4 in S_w, S_n, P_w, T).  No reference counterpart. */
    CAPAZ_ENERGY_CONFIG = 1
} capaz_energy_closure;

/* Field identifiers, in the member order of ptr_Arrays (defines.h:65-76). */
typedef enum capaz_field_id {
    CAPAZ_S_n = 0, CAPAZ_S_w, CAPAZ_P_w, CAPAZ_P_n, CAPAZ_ro_w, CAPAZ_ro_n,
    CAPAZ_ux_w, CAPAZ_uy_w, CAPAZ_uz_w, CAPAZ_ux_n, CAPAZ_uy_n, CAPAZ_uz_n,
    CAPAZ_Xi_w, CAPAZ_Xi_n, CAPAZ_roS_w, CAPAZ_roS_w_old, CAPAZ_roS_n, CAPAZ_roS_n_old,
    CAPAZ_m, CAPAZ_K,
    CAPAZ_S_g, CAPAZ_P_g, CAPAZ_ro_g, CAPAZ_ux_g, CAPAZ_uy_g, CAPAZ_uz_g, CAPAZ_Xi_g,
    CAPAZ_roS_g, CAPAZ_roS_g_old,
    CAPAZ_T, CAPAZ_H_w, CAPAZ_H_n, CAPAZ_H_g, CAPAZ_H_r, CAPAZ_E, CAPAZ_E_new,
    CAPAZ_NUM_FIELDS
} capaz_field_id;

/* Parameters of one domain: the subset of the reference's `consts` block
 * (defines.h:79-113) that the step reads, media 0 only, plus the closure
 * selector.  Field names follow the defines.ini keys in brackets. */
typedef struct capaz_params {
    unsigned struct_size;  /* sizeof(capaz_params), for ABI checks */
    int model;             /* capaz_model */
    int energy;            /* 1: energy equation (energy.cpp; two phases with
                              CAPAZ_ENERGY_CONFIG only) */
    int closure;           /* capaz_closure */
    int Nx, Ny, Nz;        /* [NX] [NY] [NZ] */
    int source;            /* [SOURCE] top-face injection half-width (>0: inject) */
    int newton_iterations; /* [ITERATIONS] */
    int device;            /* HIP device ordinal */
    double hx, hy, hz, dt; /* [HX] [HY] [HZ] [DT] */
    double P_atm, g_const; /* [P_ATM] [G_CONST] */
    double mu_w, mu_n, mu_g;    /* [MU_W] [MU_N] [MU_G] */
    double ro0_w, ro0_n, ro0_g; /* [RO_W] [RO_N] [RO_G] */
    double beta_w, beta_n;      /* [BETA_W] [BETA_N] */
    double lambda;              /* [LAMBDA_0] */
    double S_wr, S_nr, S_gr;    /* [S_WR_0] [S_NR_0] [S_GR_0] */
    double P_d;                 /* [P_D_0] two-phase entry pressure */
    double S_w_gr, S_n_gr;      /* [S_W_GR] [S_N_GR] injected saturations */
    double corey_n;             /* Corey exponent (CAPAZ_CLOSURE_COREY) */
    /* Boundary and energy closures of the benchmark configurations, no
     * reference counterpart (zero selects the reference's rule): */
    double P_bottom;            /* two phases: > 0 a fixed bottom-face pressure
                                   instead of the hydrostatic rule of
                                   two-phase.cpp:100-107 */
    int energy_closure;         /* capaz_energy_closure */
    double lambda_eff;          /* CONFIG: thermal conductivity of the medium, W/(m K) */
    double c_r, ro_r;           /* CONFIG: rock heat capacity J/(kg K), density kg/m^3 */
    double T_top, T_bottom;     /* CONFIG: top/bottom face temperatures, K;
                                   <= 0: insulated (copied from the inward cell) */
    double mu_T_a, mu_T_ref;    /* CONFIG: mu_a(T) = mu_a * exp(-mu_T_a * (T - mu_T_ref)) */
} capaz_params;

/* Fills *p with zeros, struct_size and defaults (newton_iterations = 1,
 * corey_n = 2). */
void capaz_params_init(capaz_params* p);

/* Library ABI version (CAPAZ_ABI_VERSION of the build). */
int capaz_abi_version(void);

/* device_initialization + device_memory_allocation (gpu.cu:653-852;
 * defines.h:121,138).  Allocates every field of the model on the device,
 * zero-filled, and a stream.  Grids must have Ny >= 3 and Nx, Nz each 1 or
 * >= 3 (Nx >= 3 for two phases), so that every boundary cell's inward
 * neighbour (cpu.cpp:107-141) is an interior cell. */
int capaz_create(const capaz_params* p, capaz_ctx** out);

/* device_memory_free + device_finalization (defines.h:124,139). */
int capaz_destroy(capaz_ctx* ctx);

const char* capaz_last_error(const capaz_ctx* ctx);

/* Number of cells, Nx*Ny*Nz. */
size_t capaz_cells(const capaz_ctx* ctx);

/* 1 if `field` exists for the context's model. */
int capaz_has_field(const capaz_ctx* ctx, int field);

/* load_data_to_device (gpu.cu:660-667; defines.h:135): copies n = Nx*Ny*Nz
 * doubles from host memory into a field. */
int capaz_set_field(capaz_ctx* ctx, int field, const double* host, size_t n);

/* load_data_to_host (gpu.cu:653-657; defines.h:134): copies a field to host
 * memory.  After capaz_step every field reads as it would after the same
 * number of reference time steps. */
int capaz_get_field(capaz_ctx* ctx, int field, double* host, size_t n);

/* Device address of a field's current array (valid until the next call on
 * the context), for callers that share device memory, e.g. torch tensors. */
int capaz_field_device_ptr(capaz_ctx* ctx, int field, double** dptr);

/* Device-to-device copy of n doubles into a field (stream-ordered). */
int capaz_set_field_device(capaz_ctx* ctx, int field, const double* dsrc, size_t n);

/* Per-entry-point kernels, each with the reference CPU path's semantics
 * (cpu.cpp, the model .cpp, gauss.cpp, energy.cpp), launched on the
 * context's stream without a host synchronisation. */
int capaz_S(capaz_ctx* ctx);                  /* S_calculation        cpu.cpp:28-36 */
int capaz_P_Xi_ro(capaz_ctx* ctx);            /* ro_P_Xi_calculation  cpu.cpp:3-15 */
int capaz_u(capaz_ctx* ctx);                  /* u_calculation        cpu.cpp:17-26 */
int capaz_roS(capaz_ctx* ctx, double t);      /* roS_calculation      cpu.cpp:52-65 */
int capaz_energy_HE(capaz_ctx* ctx);          /* H_E_current_calculation cpu.cpp:39-49 */
int capaz_energy_E(capaz_ctx* ctx);           /* E_calculation        cpu.cpp:68-77 */
int capaz_newton(capaz_ctx* ctx);             /* P_S_calculation      cpu.cpp:80-89 */
int capaz_boundary(capaz_ctx* ctx);           /* boundary_conditions  cpu.cpp:91-104 */

/* time_step_function (main.cpp:93-108) as the sequence of the entry-point
 * kernels above, at time t. */
int capaz_step_sequence(capaz_ctx* ctx, double t);

/* nsteps time steps, step numbers first_step .. first_step+nsteps-1
 * (t = step*dt as in main.cpp:49), with the fused kernels: the same results
 * as capaz_step_sequence, bit for bit, in fewer passes over memory. */
int capaz_step(capaz_ctx* ctx, long first_step, long nsteps);

/* Launch shape of capaz_step's kernels (does not change results):
 *   "fused_kernel"  0: one thread per cell; 1: z-marching 64 x tile_y tiles
 *                   with the plane's coefficients staged in LDS (default);
 *                   2: as 1, recomputing the centre coefficients per plane
 *   "tile_y"        0, 4 or 8 (default 0: 4 for the two-phase Corey-2 step, 8 otherwise)
 *   "kchunk"        planes per workgroup along z; 0 (default): chosen from
 *                   the grid (16 on large grids, fewer on small ones)
 *   "adjoint_kernel" capaz_adjoint's passes: 2 (default) the backward
 *                   Newton from the flux divergence recorded while each
 *                   segment is recomputed; 1 both passes on z-marching LDS
 *                   tiles; 0 both one thread per cell
 *   "adjoint_streams" 2 (default): capaz_adjoint recomputes the next
 *                   segment on a second stream while the adjoint passes
 *                   run over the current one (two segment buffers); 1: one
 *                   stream
 *   "energy_kernel" the energy model's step: 1 (default) for three phases
 *                   six passes (per cell S, P_Xi_ro, HE and the
 *                   conductivity; u; roS; E from the stored conductivity;
 *                   Newton; the boundary faces), for two phases one fused
 *                   kernel on z-marching 64 x 8 tiles (P_w, S_n, T read
 *                   once and written once per step, the Newton in
 *                   registers); 0 the eight entry-point kernels
 *   "energy_newton" the energy model's Newton (its entry point and step):
 *                   1 (default) every interior cell without the pivot
 *                   search, then the cells that need it with the search;
 *                   0 the general routine on every cell in one pass
 *   "timing"        1: record HIP events around each fused step's main
 *                   kernel (the energy model's whole step; read with
 *                   capaz_kernel_time); 0: off (default).
 *                   Setting it resets the totals.
 *   "adjoint_segment" (read only) the segment length of the last
 *                   capaz_adjoint
 *   "check_every"   K > 0: capaz_step runs capaz_check_finite after every
 *                   step whose number is a multiple of K and stops with
 *                   CAPAZ_ERR_NONFINITE at the first failing check (the
 *                   context then holds the state after that step); 0: off
 *                   (default).
 * Returns CAPAZ_ERR_INVALID for an unknown key or value.  No counterpart in
 * the reference (gpu.cu's launch shape is fixed). */
int capaz_tune(capaz_ctx* ctx, const char* key, int value);
int capaz_get_tuning(const capaz_ctx* ctx, const char* key, int* value);

/* Total device time (ms) and number of the fused step's main-kernel launches
 * timed since "timing" was last set; waits for queued work.  No counterpart
 * in the reference. */
int capaz_kernel_time(capaz_ctx* ctx, double* total_ms, long* launches);

/* Non-finite sentinel over the primary state (P_w, S_n; S_w for three
 * phases; T with the energy equation), one pass over those fields on the
 * device.  CAPAZ_OK and *first_bad = -1 when every value is finite; else
 * CAPAZ_ERR_NONFINITE, *first_bad = the smallest linear index
 * i + j*Nx + k*Nx*Ny of a cell holding a NaN or an infinity, and
 * capaz_last_error names the cell and the fields.  Replaces the printed-only
 * checks of test_nan (shared_test.cpp:56-70), which never stop a run. */
int capaz_check_finite(capaz_ctx* ctx, int64_t* first_bad);

/* The same sentinel over caller-owned device arrays, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc); it needs no context.  fields: a host array of nfields
 * (1..CAPAZ_SENTINEL_MAX_FIELDS) device pointers of n doubles each, on the
 * current device; stream: a hipStream_t (NULL: the null stream), which the
 * call waits for.  CAPAZ_OK and *first_bad = -1 when every value is finite;
 * else CAPAZ_ERR_NONFINITE and *first_bad = the smallest index holding a NaN
 * or an infinity in some field (capaz_last_error(NULL) names the fields).
 * Its only device memory is a static array of CAPAZ_SENTINEL_MAX_FIELDS
 * result slots per GPU (held until the process ends), shared by every call on
 * that GPU: calls that overlap on one GPU race; nfields above the maximum is
 * rejected with CAPAZ_ERR_INVALID. */
#define CAPAZ_SENTINEL_MAX_FIELDS 8
int capaz_check_finite_noalloc(const double* const* fields, int nfields, int64_t n, void* stream,
                               int64_t* first_bad);

/* Waits for all work queued on the context. */
int capaz_sync(capaz_ctx* ctx);

/* The context's hipStream_t, as void*. */
int capaz_stream(capaz_ctx* ctx, void** stream);

/* Face exchange buffers of the multi-domain reference (gpu.cu:854-1128;
 * defines.h:196-207): copies the second (side 0) or second-to-last (side 1)
 * plane normal to `axis` (0 x, 1 y, 2 z) of a field into `dbuf` (load), or
 * `dbuf` into the first / last plane (save).  `dbuf` is device memory of the
 * plane's size; the plane is laid out as the reference's HostBuffer. */
int capaz_load_face(capaz_ctx* ctx, int field, int axis, int side, double* dbuf);
int capaz_save_face(capaz_ctx* ctx, int field, int axis, int side, const double* dbuf);

/* One halo exchange's planes in one launch: the per-field, per-side
 * load_/save_exchange_data_part_* calls of an exchange (gpu.cu:854-1128,
 * one kernel per face there) as a single kernel.  capaz_load_faces packs,
 * for each of the nfields (1..8) fields, its second and second-to-last plane
 * normal to `axis` into dbuf as [field q][side 0, 1][plane] (each plane laid
 * out as capaz_load_face's); capaz_save_faces unpacks such a buffer into the
 * first and last planes.  dbuf: device memory of 2 * nfields planes. */
int capaz_load_faces(capaz_ctx* ctx, const int* fields, int nfields, int axis, double* dbuf);
int capaz_save_faces(capaz_ctx* ctx, const int* fields, int nfields, int axis, const double* dbuf);

/* The z exchange's pack written by the step itself (no counterpart in the
 * reference, whose exchange loads the planes after the step; for a z-slab
 * decomposition, one slab per GPU): with dbuf set, every capaz_step also
 * writes planes 1 and Nz-2 of the new state into dbuf, laid out as
 * capaz_load_faces(axis 2) of the fields P_w, S_w, S_n, P_n, P_g (three
 * phases; 10 planes) or P_w, S_n (two phases; 4 planes), in the launch that
 * writes the boundary faces (three phases) or one small launch after the step
 * (two phases).  The values are those capaz_load_faces returns after the
 * step, except P_n and P_g of the planes' interior cells, which hold the new
 * P_w (the value the next step gives them) instead of the previous one.
 * dbuf NULL turns it off.  The energy models, and contexts whose m has a zero
 * in an interior cell, are rejected (CAPAZ_ERR_INVALID, at capaz_step for
 * the latter); Nz must be at least 3.  Only capaz_step writes the pack:
 * capaz_batch_step rejects a batch with a member that has one, and the
 * history-matching calls (capaz_objective, capaz_adjoint, capaz_tangent)
 * leave it as it is. */
int capaz_set_halo_pack(capaz_ctx* ctx, double* dbuf);

/* The receiving end of that exchange, with no unpack (no counterpart in the
 * reference, whose exchange unpacks a host buffer): the device addresses of
 * the ghost planes -- plane 0 in dst[2q], plane Nz-1 in dst[2q+1] -- of the
 * exchange fields q of capaz_set_halo_pack's order, in the arrays the next
 * capaz_step reads.  A z plane of a field is Nx*Ny contiguous doubles, so a
 * neighbour's pack planes can be transferred (hipMemcpyPeerAsync, an RCCL
 * receive) straight into them.  The addresses alternate between two arrays
 * from step to step: ask again after every capaz_step.  nmax: room in dst
 * (10 for three phases, 4 for two); *n receives the count.  The same models
 * and Nz as capaz_set_halo_pack. */
int capaz_ghost_planes(capaz_ctx* ctx, double** dst, int nmax, int* n);

/* ---- Checkpoint / restart ---------------------------------------------------
 * Replace save / restore (main.cpp:1035-1145) with a format that keeps every
 * field the next step reads (the primary state, including the boundary
 * values of P_n and P_g, plus m and K; T for the energy model), so a restored
 * context continues bit for bit:
 *   "CAPAZST1", int32 version (1), model, energy, Nx, Ny, Nz, int64 step,
 *   int32 nfields, nfields int32 field ids, then nfields x Nx*Ny*Nz doubles.
 * Loading requires the same model and grid; the intermediate fields are
 * recomputed by the next step (as after capaz_set_field). */
int capaz_save_state(capaz_ctx* ctx, const char* path, long step);
int capaz_load_state(capaz_ctx* ctx, const char* path, long* step);

/* ---- Plot files ----------------------------------------------------------------
 * Appends nrows rows of a row-major nrows x ncols table to the text file at
 * path, each value printed with its column's printf format (one double
 * conversion each, e.g. "%.5e"), separated by one space, rows ending in '\n':
 * the row loop of print_plot_row (main.cpp:700-780) on host data, for
 * capaz_amd.deck's Tecplot writer.  Needs no context; errors are read with
 * capaz_last_error(NULL). */
int capaz_write_plot_rows(const char* path, const double* table, long nrows, int ncols,
                          const char* const* formats);

/* ---- History matching: objective and discrete adjoint ----------------------
 * No counterpart in the reference.  The objective of nsteps fused steps from
 * the context's current state is
 *     J = 0.5 * sum_s sum_c [(S_w(c, steps[s]) - S_w_obs[s][c])^2
 *                            + (S_n(c, steps[s]) - S_n_obs[s][c])^2]
 * over the observed cells c and steps (two phases: S_w = 1 - S_n).  The
 * three-phase model with the Corey or the reference's van Genuchten closures,
 * or the two-phase model with the Corey closure (capaz_adjoint with
 * adjoint_kernel 2), no energy equation, at most 8 Newton iterations, no
 * m == 0 cells.  capaz_tangent: three phases. */
typedef struct capaz_observations {
    size_t ncells;
    const int64_t* cells;  /* distinct linear cell indices i + j*Nx + k*Nx*Ny */
    int nsamples;
    const long* steps;     /* increasing step counts in 1..nsteps */
    const double* S_w;     /* nsamples x ncells, row-major */
    const double* S_n;     /* nsamples x ncells, row-major */
} capaz_observations;

/* J after nsteps fused steps; the context then holds the state after them
 * (as capaz_step would). */
int capaz_objective(capaz_ctx* ctx, long nsteps, const capaz_observations* obs, double* J);
/* J and its gradient with respect to ln K and m of every cell (host arrays
 * of Nx*Ny*Nz doubles, either may be NULL), by the hand-written discrete
 * adjoint of the fused step, with a checkpoint of the state every C steps and
 * each segment recomputed from its checkpoint on the way back.
 * C = checkpoint_every if > 0; round(sqrt(nsteps)) if < 0; if 0, nsteps (the
 * forward sweep records the whole trajectory and nothing is recomputed) when
 * that workspace takes at most a quarter of the device's memory and is free,
 * else round(sqrt(nsteps)).  The gradient has the same bits for every C; the
 * C used is read back with capaz_get_tuning(ctx, "adjoint_segment").  The
 * workspace stays with the context for the next call.  The context then
 * holds the state after nsteps, as after capaz_objective. */
int capaz_adjoint(capaz_ctx* ctx, long nsteps, const capaz_observations* obs, int checkpoint_every, double* J,
                  double* grad_lnK, double* grad_m);

/* J and its directional derivative dJ along (dlnK, dm) (host arrays of
 * Nx*Ny*Nz doubles, either may be NULL for zero) by the tangent-linear model
 * of the fused step, run alongside nsteps forward steps from the current
 * state (which the context then holds).  Its partials are those capaz_adjoint
 * transposes, so dJ equals the gradient's inner product with the direction
 * to rounding. */
int capaz_tangent(capaz_ctx* ctx, long nsteps, const capaz_observations* obs, const double* dlnK, const double* dm,
                  double* J, double* dJ);

/* ---- Batched ensembles ------------------------------------------------------
 * No counterpart in the reference (one domain per process there): an
 * ensemble of realizations of one physics, each with its own grid and fields,
 * advanced together -- one launch of the interior kernel and one of the
 * boundary kernel per step for all realizations, the n steps of a
 * capaz_batch_step call captured as one HIP graph and replayed.  Results are
 * those of capaz_step on each realization alone, bit for bit. */
typedef struct capaz_batch capaz_batch;

/* count realizations with the physics of `base` (its Nx/Ny/Nz ignored) and
 * grids dims[3r], dims[3r+1], dims[3r+2] (Nx, Ny, Nz of realization r).  No
 * energy model.  On failure *out is NULL and capaz_batch_last_error(NULL)
 * says why. */
int capaz_batch_create(const capaz_params* base, int count, const int* dims, capaz_batch** out);
/* Frees the batch and its realizations' contexts. */
int capaz_batch_destroy(capaz_batch* b);
const char* capaz_batch_last_error(const capaz_batch* b);
int capaz_batch_size(const capaz_batch* b);
/* The context of realization r (owned by the batch: do not destroy it).  All
 * single-domain calls work on it -- set/get fields, the entry points,
 * capaz_step -- and see the batch's steps. */
int capaz_batch_domain(capaz_batch* b, int r, capaz_ctx** ctx);
/* nsteps fused steps of every realization (step numbers from first_step, as
 * capaz_step).  Fails with CAPAZ_ERR_STATE if a realization has m == 0 cells. */
int capaz_batch_step(capaz_batch* b, long first_step, long nsteps);
/* Launch shape (results do not change):
 *   "tile_x", "tile_y"  interior tile: 64x8 (default), 64x4 or 32x8 columns
 *   "kchunk"            planes per tile (default 16)
 *   "persistent"        1: a resident grid taking tiles from a queue
 *                       (default); 0: one workgroup per tile
 *   "graph"             1: capture each run of steps as a HIP graph (default) */
int capaz_batch_tune(capaz_batch* b, const char* key, int value);
int capaz_batch_get_tuning(const capaz_batch* b, const char* key, int* value);
/* Times every tile shape x kchunk in {8,16,32} x persistent in {1,0} over
 * `steps` steps of this ensemble and keeps the fastest; the realizations'
 * state is restored afterwards.  *best_ms (may be NULL): its time per step. */
int capaz_batch_autotune(capaz_batch* b, int steps, double* best_ms);
/* Waits for the batch's and its realizations' queued work. */
int capaz_batch_sync(capaz_batch* b);
/* The batch's hipStream_t, as void*. */
int capaz_batch_stream(capaz_batch* b, void** stream);

#ifdef __cplusplus
}
#endif
#endif
