// Adjoint kernels for gfx950: the two passes of the discrete adjoint step
// (capaz_adjoint.h) and the observation gathers of the history-matching
// objective.  One thread per cell, x fastest: a 64-wide wavefront reads 64
// consecutive doubles of every field it touches, as the forward kernels do.
#include <hip/hip_runtime.h>

#include "capaz_adjoint.h"
#include "capaz_adjoint_launch.h"

namespace capaz {

namespace {

constexpr int kAdjBlock = 64;

template <int IT>
__global__ void __launch_bounds__(kAdjBlock) k_adj_newton(DevParams p, AdjIO a) {
    const int i = (int)blockIdx.x * kAdjBlock + (int)threadIdx.x, j = (int)blockIdx.y, k = (int)blockIdx.z;
    if (i >= p.Nx || !interior3(p, i, j, k)) return;
    adj_newton_cell<IT>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
}

__global__ void __launch_bounds__(kAdjBlock) k_adj_coef(DevParams p, AdjIO a) {
    const int i = (int)blockIdx.x * kAdjBlock + (int)threadIdx.x, j = (int)blockIdx.y, k = (int)blockIdx.z;
    if (i >= p.Nx) return;
    adj_coef_cell(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
}

// out_w[c] = S_w[cells[c]], out_n[c] = S_n[cells[c]]
__global__ void k_obs_gather(const double* Sw, const double* Sn, const int64_t* cells, int64_t ncells, double* out_w,
                             double* out_n) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    out_w[c] = Sw[cells[c]];
    out_n[c] = Sn[cells[c]];
}

// lam_S[cells[c]] += seen[c] - want[c] (cells distinct)
__global__ void k_obs_adjoint(double* lSw, double* lSn, const int64_t* cells, int64_t ncells, const double* seen_w,
                              const double* seen_n, const double* want_w, const double* want_n) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    const int64_t l = cells[c];
    lSw[l] += seen_w[c] - want_w[c];
    lSn[l] += seen_n[c] - want_n[c];
}

}  // namespace

hipError_t launch_adjoint_step(const DevParams& p, const AdjIO& a, hipStream_t s) {
    const dim3 grid((unsigned)((p.Nx + kAdjBlock - 1) / kAdjBlock), (unsigned)p.Ny, (unsigned)p.Nz);
    // the common iteration counts keep the Newton iterates in registers
    switch (p.iterations) {
        case 1: hipLaunchKernelGGL(k_adj_newton<1>, grid, dim3(kAdjBlock), 0, s, p, a); break;
        case 2: hipLaunchKernelGGL(k_adj_newton<2>, grid, dim3(kAdjBlock), 0, s, p, a); break;
        case 3: hipLaunchKernelGGL(k_adj_newton<3>, grid, dim3(kAdjBlock), 0, s, p, a); break;
        case 4: hipLaunchKernelGGL(k_adj_newton<4>, grid, dim3(kAdjBlock), 0, s, p, a); break;
        default: hipLaunchKernelGGL(k_adj_newton<0>, grid, dim3(kAdjBlock), 0, s, p, a); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_adj_coef, grid, dim3(kAdjBlock), 0, s, p, a);
    return hipGetLastError();
}

hipError_t launch_obs_gather(const double* Sw, const double* Sn, const int64_t* cells, int64_t ncells, double* out_w,
                             double* out_n, hipStream_t s) {
    if (ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_obs_gather, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, s, Sw, Sn, cells, ncells,
                       out_w, out_n);
    return hipGetLastError();
}

hipError_t launch_obs_adjoint(double* lSw, double* lSn, const int64_t* cells, int64_t ncells, const double* seen_w,
                              const double* seen_n, const double* want_w, const double* want_n, hipStream_t s) {
    if (ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_obs_adjoint, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, s, lSw, lSn, cells, ncells,
                       seen_w, seen_n, want_w, want_n);
    return hipGetLastError();
}

}  // namespace capaz
