// Launchers of the backend's kernels (defined in capaz_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "capaz_bodies.h"

namespace capaz {

enum class Entry { S, P_Xi_ro, U, RoS, HE, E, Newton, Boundary };

hipError_t launch_entry(Entry e, const DevParams& p, const DevFields& F, hipStream_t s);
hipError_t launch_fused_step(const DevParams& p, const StepIO& io, hipStream_t s);
hipError_t launch_copy_interior(const DevParams& p, const double* src, double* dst, hipStream_t s);
hipError_t launch_m_zero(const DevParams& p, const double* m, int* flag, hipStream_t s);
hipError_t launch_face(const DevParams& p, double* field, double* buf, int axis, int pos, int save, hipStream_t s);

}  // namespace capaz
