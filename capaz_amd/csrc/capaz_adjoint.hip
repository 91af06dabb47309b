// Adjoint kernels for gfx950: the two passes of the discrete adjoint step
// (capaz_adjoint.h) and the observation gathers of the history-matching
// objective.  One thread per cell, x fastest: a 64-wide wavefront reads 64
// consecutive doubles of every field it touches, as the forward kernels do.
#include <hip/hip_runtime.h>

#include "capaz_adjoint.h"
#include "capaz_adjoint_launch.h"
#include "capaz_fused.h"

namespace capaz {

namespace {

constexpr int kAdjBlock = 64;

template <int IT>
__global__ void __launch_bounds__(kAdjBlock) k_adj_newton(DevParams p, AdjIO a) {
    const int i = (int)blockIdx.x * kAdjBlock + (int)threadIdx.x, j = (int)blockIdx.y, k = (int)blockIdx.z;
    if (i >= p.Nx || !interior3(p, i, j, k)) return;
    adj_newton_cell<IT>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
}

// Interior cells, densely numbered (every lane of a wavefront has a cell).
template <int IT>
__global__ void __launch_bounds__(256) k_adj_newton_rec(DevParams p, AdjIO a, const double* div, int64_t stride,
                                                        uint32_t ni, uint32_t nj, uint32_t total) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= total) return;
    int i, j, k;
    box_cell(t, ni, nj, p.Nx >= 3 ? 1 : 0, 1, p.Nz >= 3 ? 1 : 0, i, j, k);
    adj_newton_recorded<IT>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz, div, stride);
}

// Every cell, in memory order.  Four waves per SIMD (128 VGPRs; the
// compiler's own choice of 130 allows three) hide more of the neighbours'
// load latency, at the cost of a 12-byte spill.
// The grid is a multiple of the 8 XCDs; each XCD takes a contiguous run of
// cells (workgroups go to the XCDs round robin), so the z neighbours a cell
// gathers from were read by the same XCD and hit its L2.
// CL: the closure (1 Corey, 0 the reference's van Genuchten).
template <int CL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_adj_coef(DevParams p, AdjIO a) {
    const uint32_t b = (blockIdx.x % 8u) * (gridDim.x / 8u) + blockIdx.x / 8u;
    const uint32_t t = b * 256u + threadIdx.x;
    if (t >= (uint32_t)p.n) return;
    int i, j, k;
    box_cell(t, (uint32_t)p.Nx, (uint32_t)p.Ny, 0, 0, 0, i, j, k);
    adj_coef_cell<CL>(p, a, i, j, k, (int64_t)t);
}

// Two phases: pass 1 over the interior cells (dense numbering), pass 2 over
// every cell (XCD-contiguous runs, as k_adj_coef).
template <int IT>
__global__ void __launch_bounds__(256) k_adj2_newton_rec(DevParams p, AdjIO a, const double* div, int64_t stride,
                                                         uint32_t ni, uint32_t nj, uint32_t total) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= total) return;
    int i, j, k;
    box_cell(t, ni, nj, 1, 1, p.Nz >= 3 ? 1 : 0, i, j, k);
    adj2_newton_recorded<IT>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz, div, stride);
}

template <bool N2>
__global__ void __launch_bounds__(256) k_adj2_coef(DevParams p, AdjIO a) {
    const uint32_t b = (blockIdx.x % 8u) * (gridDim.x / 8u) + blockIdx.x / 8u;
    const uint32_t t = b * 256u + threadIdx.x;
    if (t >= (uint32_t)p.n) return;
    int i, j, k;
    box_cell(t, (uint32_t)p.Nx, (uint32_t)p.Ny, 0, 0, 0, i, j, k);
    adj2_coef_cell<N2>(p, a, i, j, k, (int64_t)t);
}

// Tangent-linear step: interior cells (dense numbering), then the faces.
__global__ void __launch_bounds__(256) k_tan_interior(DevParams p, TanIO a, uint32_t ni, uint32_t nj, uint32_t total) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    if (t >= total) return;
    int i, j, k;
    box_cell(t, ni, nj, p.Nx >= 3 ? 1 : 0, 1, p.Nz >= 3 ? 1 : 0, i, j, k);
    tan_cell<0>(p, a, i, j, k, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
}

__global__ void __launch_bounds__(256) k_tan_faces(DevParams p, TanIO a) {
    int i, j, k;
    if (!face_cell(p, (int64_t)blockIdx.x * 256 + threadIdx.x, i, j, k)) return;
    tan_boundary(p, a, i, j, k);
}

// out_w[c] = S_w[cells[c]], out_n[c] = S_n[cells[c]]; Sw == NULL (two
// phases): S_w = 1 - S_n.
__global__ void k_obs_gather(const double* Sw, const double* Sn, const int64_t* cells, int64_t ncells, double* out_w,
                             double* out_n) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    const double sn = Sn[cells[c]];
    out_w[c] = Sw ? Sw[cells[c]] : 1. - sn;
    out_n[c] = sn;
}

// lam_S[cells[c]] += seen[c] - want[c] (cells distinct); lSw == NULL (two
// phases, S_w = 1 - S_n): the S_w mismatch goes to lam_S_n with a minus sign.
__global__ void k_obs_adjoint(double* lSw, double* lSn, const int64_t* cells, int64_t ncells, const double* seen_w,
                              const double* seen_n, const double* want_w, const double* want_n) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= ncells) return;
    const int64_t l = cells[c];
    if (lSw) {
        lSw[l] += seen_w[c] - want_w[c];
        lSn[l] += seen_n[c] - want_n[c];
    } else {
        lSn[l] += (seen_n[c] - want_n[c]) - (seen_w[c] - want_w[c]);
    }
}

// ---- tiled passes: the z-marching tiles of the fused step -------------------

// Pass 1 as the fused step's epilogue: the tile computes the cell's
// coefficients and flux divergence exactly as the step does, then runs the
// Newton iteration backwards instead of forwards.
template <int IT>
struct AdjNewtonStore {
    AdjIO a;
    __device__ __forceinline__ void operator()(const DevParams& p, int i, int j, int k, int64_t l, const CoefT<3>& c,
                                               const double* div) const {
        adj_newton_tail<IT>(p, a, i, j, k, l, c.ro, c.P[0], c.Sw, c.Sn, div);
    }
};

constexpr int kTX = 64, kTY = 8;

template <int CLOSURE, int IT>
__global__ void __launch_bounds__(kTX * kTY) k_adj_newton_z(DevParams p, StepIO io, AdjIO a, int kchunk) {
    __shared__ FusedLds<9, kTX, kTY> lds;
    fused_z_tile<3, CLOSURE, kTX, kTY, true>(p, io, kchunk, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, lds,
                                             AdjNewtonStore<IT>{a});
}

// Pass 2 over tiles of kTX x kTY columns (every cell, boundary included)
// marching kchunk planes: each plane's coefficients and mu, with a one-cell
// halo ring, are staged in LDS once and read by the x / y neighbours; the
// z neighbours are carried in registers.
struct AdjCoefLds {
    double v[12][kTY + 2][kTX + 2];  // P[3], Xi[3], ro[3], mu[3]
};

template <int CL>
__device__ __forceinline__ AdjNb adj_stage(const DevParams& p, const AdjIO& a, int i, int j, int k, bool in_grid) {
    AdjNb n;
    n.exists = in_grid;
    n.interior = false;
    n.mu[0] = n.mu[1] = n.mu[2] = 0.;
    n.c = Coef{};
    if (!in_grid) return n;
    const int64_t q = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz;
    n.interior = interior3(p, i, j, k);
    n.c = coef_v<CL>(p, a, q, n.interior);
    if (n.interior) {
        n.mu[0] = a.mu_w[q];
        n.mu[1] = a.mu_n[q];
        n.mu[2] = a.mu_g[q];
    }
    return n;
}

__device__ __forceinline__ void lds_put(AdjCoefLds& lds, int y, int x, const AdjNb& n) {
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        lds.v[ph][y][x] = n.c.P[ph];
        lds.v[3 + ph][y][x] = n.c.Xi[ph];
        lds.v[6 + ph][y][x] = n.c.ro[ph];
        lds.v[9 + ph][y][x] = n.mu[ph];
    }
}

__device__ __forceinline__ AdjNb lds_get(const DevParams& p, const AdjCoefLds& lds, int y, int x, int i, int j,
                                         int k) {
    AdjNb n;
    n.exists = i >= 0 && i < p.Nx && j >= 0 && j < p.Ny;
    n.interior = n.exists && interior3(p, i, j, k);
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        n.c.P[ph] = lds.v[ph][y][x];
        n.c.Xi[ph] = lds.v[3 + ph][y][x];
        n.c.ro[ph] = lds.v[6 + ph][y][x];
        n.mu[ph] = lds.v[9 + ph][y][x];
    }
    return n;
}

template <int CL>
__global__ void __launch_bounds__(kTX * kTY) k_adj_coef_z(DevParams p, AdjIO a, int kchunk) {
    __shared__ AdjCoefLds lds;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = kTX * (int)blockIdx.x, j0 = kTY * (int)blockIdx.y, k0 = kchunk * (int)blockIdx.z;
    const int k1 = min(p.Nz - 1, k0 + kchunk - 1);
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < p.Nx && j < p.Ny;
    // halo cell of this thread, as in the fused step's tiles
    int hi = -1, hj = -1, hx = 0, hy = 0;
    if (ty == 0) { hi = i; hj = j0 - 1; hx = tx + 1; hy = 0; }
    else if (ty == 1) { hi = i; hj = j0 + kTY; hx = tx + 1; hy = kTY + 1; }
    else if (ty == 2 && tx < kTY) { hi = i0 - 1; hj = j0 + tx; hx = 0; hy = tx + 1; }
    else if (ty == 2 && tx >= kTY && tx < 2 * kTY) { hi = i0 + kTX; hj = j0 + tx - kTY; hx = kTX + 1; hy = tx - kTY + 1; }
    const bool halo = hi >= 0 && hi < p.Nx && hj >= 0 && hj < p.Ny;
    const bool hasz = !(p.Nz < 2);

    AdjNb zm = adj_stage<CL>(p, a, i, j, k0 - 1, own && hasz && k0 >= 1);
    AdjNb c = adj_stage<CL>(p, a, i, j, k0, own);
    for (int k = k0; k <= k1; ++k) {
        if (own) lds_put(lds, ty + 1, tx + 1, c);
        if (halo) lds_put(lds, hy, hx, adj_stage<CL>(p, a, hi, hj, k, true));
        __syncthreads();
        const AdjNb zp = adj_stage<CL>(p, a, i, j, k + 1, own && hasz && k + 1 < p.Nz);
        if (own) {
            const int64_t l = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz;
            const CoefD cd = coef_d<CL>(p, a, l, c.interior);
            double lP[3] = {0., 0., 0.}, lL[3] = {0., 0., 0.}, lro[3] = {0., 0., 0.};
            if (!(p.Nx < 2))
                adj_coef_axis(p, cd, c.mu, lds_get(p, lds, ty + 1, tx, i - 1, j, k),
                              lds_get(p, lds, ty + 1, tx + 2, i + 1, j, k), p.r_hx, false, lP, lL, lro);
            adj_coef_axis(p, cd, c.mu, lds_get(p, lds, ty, tx + 1, i, j - 1, k),
                          lds_get(p, lds, ty + 2, tx + 1, i, j + 1, k), p.r_hy, true, lP, lL, lro);
            if (hasz) adj_coef_axis(p, cd, c.mu, zm, zp, p.r_hz, false, lP, lL, lro);
            adj_coef_finish(p, a, l, cd, lP, lL, lro);
        }
        __syncthreads();
        zm = c;
        c = zp;
    }
}

hipError_t launch_adj_coef_z(const DevParams& p, const AdjIO& a, dim3 g, int kchunk, hipStream_t s) {
    if (p.closure == 0)
        hipLaunchKernelGGL(k_adj_coef_z<0>, g, dim3(kTX, kTY), 0, s, p, a, kchunk);
    else if (p.corey_n == 2.)
        hipLaunchKernelGGL(k_adj_coef_z<2>, g, dim3(kTX, kTY), 0, s, p, a, kchunk);
    else
        hipLaunchKernelGGL(k_adj_coef_z<1>, g, dim3(kTX, kTY), 0, s, p, a, kchunk);
    return hipGetLastError();
}

// Planes per tile: enough workgroups to fill the GPU several times over,
// at most 16 planes (the fused step's default) and at least 2.
int adj_kchunk(int nz, int64_t tiles_xy) {
    int64_t kc = (nz * tiles_xy + 2047) / 2048;
    return (int)(kc < 2 ? 2 : (kc > 16 ? 16 : kc));
}

}  // namespace

hipError_t launch_adjoint_step_tiled(const DevParams& p, const AdjIO& a, hipStream_t s) {
    StepIO io{};
    io.Pw = a.Pw;
    io.Sw = a.Sw;
    io.Sn = a.Sn;
    io.Pn = a.Pn;
    io.Pg = a.Pg;
    io.m = a.m;
    io.K = a.K;
    io.dtm = a.dtm;
    const int nx = p.Nx >= 3 ? p.Nx - 2 : 1, ny = p.Ny - 2, nz = p.Nz >= 3 ? p.Nz - 2 : 1;
    const int64_t txy = (int64_t)((nx + kTX - 1) / kTX) * ((ny + kTY - 1) / kTY);
    const int kc1 = adj_kchunk(nz, txy);
    const dim3 g1((unsigned)((nx + kTX - 1) / kTX), (unsigned)((ny + kTY - 1) / kTY), (unsigned)((nz + kc1 - 1) / kc1));
    const int closure = p.closure == 0 ? 0 : (p.corey_n == 2. ? 2 : 1);
#define CAPAZ_ADJ1(IT)                                                                                         \
    do {                                                                                                       \
        if (closure == 2)                                                                                      \
            hipLaunchKernelGGL((k_adj_newton_z<2, IT>), g1, dim3(kTX, kTY), 0, s, p, io, a, kc1);             \
        else if (closure == 1)                                                                                 \
            hipLaunchKernelGGL((k_adj_newton_z<1, IT>), g1, dim3(kTX, kTY), 0, s, p, io, a, kc1);             \
        else                                                                                                   \
            hipLaunchKernelGGL((k_adj_newton_z<0, IT>), g1, dim3(kTX, kTY), 0, s, p, io, a, kc1);             \
    } while (0)
    switch (p.iterations) {
        case 1: CAPAZ_ADJ1(1); break;
        case 2: CAPAZ_ADJ1(2); break;
        case 3: CAPAZ_ADJ1(3); break;
        case 4: CAPAZ_ADJ1(4); break;
        default: CAPAZ_ADJ1(0); break;
    }
#undef CAPAZ_ADJ1
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t txy2 = (int64_t)((p.Nx + kTX - 1) / kTX) * ((p.Ny + kTY - 1) / kTY);
    const int kc2 = adj_kchunk(p.Nz, txy2);
    const dim3 g2((unsigned)((p.Nx + kTX - 1) / kTX), (unsigned)((p.Ny + kTY - 1) / kTY),
                  (unsigned)((p.Nz + kc2 - 1) / kc2));
    return launch_adj_coef_z(p, a, g2, kc2, s);
}

hipError_t launch_adjoint_coef(const DevParams& p, const AdjIO& a, hipStream_t s) {
    const dim3 grid((unsigned)((p.n + 255) / 256 + 7) / 8 * 8);
    if (p.closure == 0)
        hipLaunchKernelGGL(k_adj_coef<0>, grid, dim3(256), 0, s, p, a);
    else if (p.corey_n == 2.)
        hipLaunchKernelGGL(k_adj_coef<2>, grid, dim3(256), 0, s, p, a);
    else
        hipLaunchKernelGGL(k_adj_coef<1>, grid, dim3(256), 0, s, p, a);
    return hipGetLastError();
}

hipError_t launch_adjoint_step_recorded(const DevParams& p, const AdjIO& a, const double* div, int64_t stride,
                                        hipStream_t s) {
    const uint32_t ni = p.Nx >= 3 ? p.Nx - 2 : 1, nj = p.Ny - 2, nk = p.Nz >= 3 ? p.Nz - 2 : 1;
    const uint32_t total = ni * nj * nk;
    if (p.model == 2) {
        hipLaunchKernelGGL(k_adj2_newton_rec<0>, dim3((total + 255) / 256), dim3(256), 0, s, p, a, div, stride, ni, nj,
                           total);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const dim3 grid((unsigned)((p.n + 255) / 256 + 7) / 8 * 8);
        if (p.corey_n == 2.)
            hipLaunchKernelGGL(k_adj2_coef<true>, grid, dim3(256), 0, s, p, a);
        else
            hipLaunchKernelGGL(k_adj2_coef<false>, grid, dim3(256), 0, s, p, a);
        return hipGetLastError();
    }
    // the run-time iteration count: its loop holds fewer live registers
    // (144 VGPRs) than the fully unrolled fixed counts, without scratch
    hipLaunchKernelGGL(k_adj_newton_rec<0>, dim3((total + 255) / 256), dim3(256), 0, s, p, a, div, stride, ni, nj,
                       total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_adjoint_coef(p, a, s);
}

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
    return launch_adjoint_coef(p, a, s);
}

hipError_t launch_tangent_step(const DevParams& p, const TanIO& a, hipStream_t s) {
    const uint32_t ni = p.Nx >= 3 ? p.Nx - 2 : 1, nj = p.Ny - 2, nk = p.Nz >= 3 ? p.Nz - 2 : 1;
    const uint32_t total = ni * nj * nk;
    hipLaunchKernelGGL(k_tan_interior, dim3((total + 255) / 256), dim3(256), 0, s, p, a, ni, nj, total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nf = fused_face_cells(p);
    hipLaunchKernelGGL(k_tan_faces, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, p, a);
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
