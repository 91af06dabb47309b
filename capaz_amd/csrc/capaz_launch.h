// Launchers of the backend's kernels (defined in capaz_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "capaz_bodies.h"

namespace capaz {

enum class Entry { S, P_Xi_ro, U, RoS, HE, E, Newton, Boundary };

// Launch shape of the fused step: fused_kernel 0 = one thread per cell
// (k_fused_step), 1 = z-marching tiles (k_fused_z) of 64 x tile_y columns
// and kchunk planes per workgroup (0: chosen from the grid) that carry the
// next plane's coefficients in registers, 2 = the same tiles recomputing
// them per plane.
// adjoint_kernel: 2 = the backward Newton per cell from the flux divergence
// recorded by the segment's forward recompute, then the coefficient pass per
// cell; 1 = both passes on the z-marching LDS tiles; 0 = both per cell,
// recomputing the neighbours' coefficients.
// adjoint_streams: 2 = the reverse sweep recomputes segment q-1 on a second
// stream while the adjoint passes run over segment q (two segment buffers),
// 1 = one stream, recompute then passes.
// energy_kernel: 1 = the three-phase energy model's step in six passes (S,
// P_Xi_ro, HE and lambda_eff per cell; boundary u; roS and interior u on z
// tiles; E on z tiles from the stored lambda_eff; Newton; boundary faces),
// the two-phase energy model's in one fused kernel; 0 = the eight
// entry-point kernels.
// energy_newton: 1 = the energy model's Newton in two passes (every interior
// cell without the pivot search, then the cells that need it), 0 = the
// general routine in one pass; the results are the same.
struct Tuning {
    int check_every = 0;  // non-finite sentinel after every check_every-th step (0: off)
    int fused_kernel = 1;
    int tile_y = 0;  // 0: automatic (4 for the two-phase Corey-2 step, else 8)
    int kchunk = 0;
    int adjoint_kernel = 2;
    int adjoint_streams = 2;
    int energy_kernel = 1;
    int energy_newton = 1;
};

// Optional events recorded around the fused step's main kernel (the
// z-marching tile kernel, or the one-thread-per-cell kernel), not its
// boundary-face kernel.
struct KernelEvents {
    hipEvent_t begin, end;
};

hipError_t launch_entry(Entry e, const DevParams& p, const DevFields& F, hipStream_t s);
hipError_t launch_fused_step(const DevParams& p, const StepIO& io, hipStream_t s);
hipError_t launch_fused(const DevParams& p, const StepIO& io, const Tuning& t, hipStream_t s,
                        const KernelEvents* ev = nullptr);
// The roS entry point of the three-phase model on the fused kernel's z tiles,
// from the P, Xi and ro arrays P_Xi_ro wrote (no cell may have m == 0: the
// per-cell entry point keeps their stale roS); dtm holds dt / m per cell.
hipError_t launch_roS_stored(const DevParams& p, const DevFields& F, const double* dtm, hipStream_t s);
// E_new of the interior cells on z tiles (energy model); lam: lambda_eff per
// cell computed earlier in the step, or NULL to evaluate it from the fields.
hipError_t launch_energy_e_tiles(const DevParams& p, const DevFields& F, const double* lam, hipStream_t s);
// One step of the energy model in six passes, in place on F; lam is a work
// array of Nx*Ny*Nz doubles, dtm holds dt / m per cell, redo Nx*Ny*Nz bytes
// (rounded up to 16) for the Newton pass (NULL: the Newton in one pass, as
// launch_newton_energy); ev brackets the whole step.
hipError_t launch_energy_step(const DevParams& p, const DevFields& F, double* lam, const double* dtm,
                              unsigned char* redo, hipStream_t s, const KernelEvents* ev = nullptr);
// The energy model's Newton entry point: interior cells without the pivot
// search, then the cells that need it (marked in redo, Nx*Ny*Nz bytes rounded
// up to 16) with the general routine; redo NULL: the general routine on every
// cell in one pass.
hipError_t launch_newton_energy(const DevParams& p, const DevFields& F, unsigned char* redo, hipStream_t s);
// One fused step of the two-phase energy model (configuration closure):
// cur = {P_w, S_n, T, m, K}, next = {P_w, S_n, T} (every cell written);
// kchunk as for the fused step; ev brackets the kernel.
hipError_t launch_energy2_step(const DevParams& p, const double* const* cur, double* const* next, int kchunk,
                               hipStream_t s, const KernelEvents* ev = nullptr);
// Planes per tile of the fused z kernels: kchunk if >= 1, else chosen from
// the grid and the number of workgroups resident at once, so that small grids
// fill the GPU without a nearly empty last round.
int fused_kchunk(const DevParams& p, int kchunk, int ty, int resident);
// One fused three-phase step (Corey closures, 64 x 8 tiles) that also writes
// each interior cell's flux divergence of phase a to div[a*stride + l].
hipError_t launch_fused_record(const DevParams& p, const StepIO& io, int kchunk, double* div, int64_t stride,
                               hipStream_t s);
hipError_t launch_dtm(const DevParams& p, const double* m, double* dtm, hipStream_t s);
hipError_t launch_copy_interior(const DevParams& p, const double* src, double* dst, hipStream_t s);
hipError_t launch_m_zero(const DevParams& p, const double* m, int* flag, hipStream_t s);
// One face plane; a load with `inner` set reads the plane's interior cells
// from it instead of the field.
hipError_t launch_face(const DevParams& p, double* field, const double* inner, double* buf, int axis, int pos,
                       int save, hipStream_t s);
// Both exchange planes of nfields (<= kMaxFaceFields) fields along `axis` in
// one launch, buffer layout [field][side][plane]; inner (NULL, or NULL or a
// source per field) as launch_face's.
constexpr int kMaxFaceFields = 8;
hipError_t launch_faces(const DevParams& p, double* const* fields, const double* const* inner, int nfields,
                        double* buf, int axis, int save, hipStream_t s);
// *first = the smallest cell index at which f[0], f[1] or (where not NULL)
// f[2], f[3] is not finite, ~0 if there is none.
hipError_t launch_find_nonfinite(const DevParams& p, const double* const* f, unsigned long long* first,
                                 hipStream_t s);

}  // namespace capaz
