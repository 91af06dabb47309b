// Launchers of the backend's kernels (defined in capaz_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "capaz_bodies.h"

namespace capaz {

enum class Entry { S, P_Xi_ro, U, RoS, HE, E, Newton, Boundary };

// Launch shape of the fused step: fused_kernel 0 = one thread per cell
// (k_fused_step), 1 = z-marching tiles (k_fused_z) of 64 x tile_y columns
// and kchunk planes per workgroup that carry the next plane's coefficients
// in registers, 2 = the same tiles recomputing them per plane.
// adjoint_kernel: 1 = the adjoint passes on the same z-marching LDS tiles,
// 0 = one thread per cell reading neighbours from memory.
struct Tuning {
    int fused_kernel = 1;
    int tile_y = 8;
    int kchunk = 16;
    int adjoint_kernel = 1;
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
hipError_t launch_dtm(const DevParams& p, const double* m, double* dtm, hipStream_t s);
hipError_t launch_copy_interior(const DevParams& p, const double* src, double* dst, hipStream_t s);
hipError_t launch_m_zero(const DevParams& p, const double* m, int* flag, hipStream_t s);
hipError_t launch_face(const DevParams& p, double* field, double* buf, int axis, int pos, int save, hipStream_t s);

}  // namespace capaz
