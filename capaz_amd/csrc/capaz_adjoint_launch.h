// Launchers of the adjoint kernels (capaz_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "capaz_adjoint.h"

namespace capaz {

// Both passes of one adjoint step over the whole grid: one thread per cell
// reading neighbours from memory, or the z-marching LDS tiles.
hipError_t launch_adjoint_step(const DevParams& p, const AdjIO& a, hipStream_t s);
hipError_t launch_adjoint_step_tiled(const DevParams& p, const AdjIO& a, hipStream_t s);
// The same with pass 1 reading the flux divergence recorded by the forward
// step (div_a at div + a*stride) instead of recomputing the stencil.
hipError_t launch_adjoint_step_recorded(const DevParams& p, const AdjIO& a, const double* div, int64_t stride,
                                        hipStream_t s);
// One tangent-linear step (interior cells, then the faces).
hipError_t launch_tangent_step(const DevParams& p, const TanIO& a, hipStream_t s);
hipError_t launch_obs_gather(const double* Sw, const double* Sn, const int64_t* cells, int64_t ncells, double* out_w,
                             double* out_n, hipStream_t s);
hipError_t launch_obs_adjoint(double* lSw, double* lSn, const int64_t* cells, int64_t ncells, const double* seen_w,
                              const double* seen_n, const double* want_w, const double* want_n, hipStream_t s);

}  // namespace capaz
