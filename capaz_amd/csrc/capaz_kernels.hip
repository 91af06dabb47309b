// HIP kernels of the backend for gfx950 and their launchers.  The per-cell
// arithmetic lives in capaz_bodies.h / capaz_cell.h; the kernels here map
// threads onto cells (x fastest, so a wavefront reads 64 consecutive doubles
// of each field) and are launched on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "capaz_bodies.h"
#include "capaz_fused.h"
#include "capaz_launch.h"

namespace capaz {

namespace {

constexpr int kBlock = 256;

__global__ void __launch_bounds__(kBlock) k_S(DevParams p, DevFields F) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_S(p, F, l);
}
__global__ void __launch_bounds__(kBlock) k_P_Xi_ro(DevParams p, DevFields F) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_P_Xi_ro(p, F, l);
}
__global__ void __launch_bounds__(kBlock) k_u(DevParams p, DevFields F) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_u(p, F, l);
}
__global__ void __launch_bounds__(kBlock) k_roS(DevParams p, DevFields F) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_roS(p, F, l);
}
__global__ void __launch_bounds__(kBlock) k_HE(DevParams p, DevFields F) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_HE(p, F, l);
}
__global__ void __launch_bounds__(kBlock) k_E(DevParams p, DevFields F) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_E(p, F, l);
}
__global__ void __launch_bounds__(kBlock) k_newton(DevParams p, DevFields F) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_newton(p, F, l);
}
__global__ void __launch_bounds__(kBlock) k_fused_step(DevParams p, StepIO io) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_fused_step(p, io, l);
}

// Energy step passes of their own: A and u of the boundary cells (face_cell's
// numbering); roS with the interior u, and E, run on z tiles (capaz_fused.h);
// Newton and the boundary are the entry points.
__global__ void __launch_bounds__(kBlock) k_energy_a(DevParams p, DevFields F, double* lam) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) body_energy_a(p, F, lam, l);
}
__global__ void __launch_bounds__(kBlock) k_energy_u_faces(DevParams p, DevFields F) {
    int i, j, k;
    if (face_cell(p, (int64_t)blockIdx.x * kBlock + threadIdx.x, i, j, k))
        body_u(p, F, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
}
// The energy model's Newton update: every interior cell without the pivot
// search (166 VGPRs, no scratch: 3 waves per SIMD against 2 for the general
// routine); a cell whose matrices need the search is marked in `redo` and
// recomputed from its untouched state by k_newton_energy_redo.
__global__ void __launch_bounds__(kBlock) k_newton_energy(DevParams p, DevFields F, unsigned char* redo) {
    const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l >= p.n) return;
    int i, j, k;
    decompose(p, l, i, j, k);
    if (p.model == 2)
        redo[l] = interior_newton2(p, i, j, k) && !body_newton_energy2_nopivot(p, F, l) ? 1 : 0;
    else
        redo[l] = interior3(p, i, j, k) && !body_newton_energy_nopivot(p, F, l) ? 1 : 0;
}
// One thread per 16 marks (one 16-byte load; redo is allocated rounded up to
// 16 bytes, the marks past Nx*Ny*Nz are never read as cells).
__global__ void __launch_bounds__(kBlock) k_newton_energy_redo(DevParams p, DevFields F, const unsigned char* redo) {
    const int64_t l0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 16;
    if (l0 >= p.n) return;
    const uint4 w = *reinterpret_cast<const uint4*>(redo + l0);
    if ((w.x | w.y | w.z | w.w) == 0u) return;
    const int64_t end = l0 + 16 < p.n ? l0 + 16 : p.n;
    for (int64_t l = l0; l < end; ++l)
        if (redo[l]) body_newton(p, F, l);
}

// boundary_conditions over the boundary cells only (the entry point and the
// energy step's last pass)
__global__ void __launch_bounds__(kBlock) k_boundary_faces(DevParams p, DevFields F) {
    int i, j, k;
    if (face_cell(p, (int64_t)blockIdx.x * kBlock + threadIdx.x, i, j, k))
        body_boundary(p, F, (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz);
}

// Copies the interior cells of src into dst (the three-phase P_n/P_g whose
// interior values belong to the previous state after a fused step).
__global__ void __launch_bounds__(kBlock) k_copy_interior(DevParams p, const double* src, double* dst) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l >= p.n) return;
    int i, j, k;
    decompose(p, l, i, j, k);
    if (interior3(p, i, j, k)) dst[l] = src[l];
}

// dt / m of every cell, the factor of the mass update (cpu.cpp:707-708, 723),
// computed once when the porosity changes instead of once per step.
__global__ void __launch_bounds__(kBlock) k_dtm(DevParams p, const double* m, double* dtm) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l < p.n) dtm[l] = p.dt / m[l];
}

// Porosity check of the fused path: sets *flag if an interior cell has m == 0
// (assign_roS_nr skips such cells, cpu.cpp:625-626).
__global__ void __launch_bounds__(kBlock) k_m_zero(DevParams p, const double* m, int* flag) {
    int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (l >= p.n) return;
    int i, j, k;
    decompose(p, l, i, j, k);
    if (cell_inner(p, i, j, k) && !m[l]) atomicOr(flag, 1);
}

// field[l], or inner[l] where inner is set and the cell is interior.  Both
// loads are made under a uniform test of inner and the choice is a per-lane
// select of values: a conditional choice between the two base pointers was
// compiled into a uniform pointer (every lane reading inner).
__device__ __forceinline__ double pick_inner(const double* field, const double* inner, int64_t l, bool interior) {
    double v = field[l];
    if (inner) {
        const double w = inner[l];
        v = interior ? w : v;
    }
    return v;
}

// Face planes (cpu.cpp:769-888): plane index `pos` normal to `axis`; the
// buffer is indexed (a + na*b) with (a, b) = (j, k) for x, (i, k) for y,
// (i, j) for z.  A load with `inner` set takes the plane's interior cells
// (interior3) from `inner` instead of the field: P_n and P_g of a
// three-phase state after fused steps, whose interior values are the
// previous P_w (what the last step's P_Xi_ro gave them; the fused step
// keeps them in registers and stores only their boundary values).
__global__ void __launch_bounds__(kBlock) k_face(DevParams p, double* field, const double* inner, double* buf,
                                                 int axis, int pos, int save) {
    int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int na = axis == 0 ? p.Ny : p.Nx;
    int nb = axis == 2 ? p.Ny : p.Nz;
    if (t >= (int64_t)na * nb) return;
    int a = (int)(t % na), b = (int)(t / na);
    int i, j, k;
    if (axis == 0) { i = pos; j = a; k = b; }
    else if (axis == 1) { i = a; j = pos; k = b; }
    else { i = a; j = b; k = pos; }
    int64_t l = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz;
    if (save) {
        field[l] = buf[t];
        return;
    }
    buf[t] = pick_inner(field, inner, l, interior3(p, i, j, k));
}

// All the exchange planes of one halo exchange in one launch.  Grid row y
// is side s of field q (y = q * 2 + s, the buffer layout: that side's plane
// at buf + y * plane, cell c = a + na * b as k_face numbers it), so the
// field pointer and the plane position are uniform over a workgroup.  A z
// plane is contiguous in the field too (sy == Nx); with vec set (checked on
// the host: even plane, 16-byte aligned field and buffer) each lane moves two
// doubles of it.  inner[q]: as k_face's `inner`, for field q of a load.
struct FaceFields {
    double* f[kMaxFaceFields];
    const double* inner[kMaxFaceFields];
};
__global__ void __launch_bounds__(kBlock) k_faces(DevParams p, FaceFields ff, double* buf, int axis, int save,
                                                  int vec) {
    const int y = (int)blockIdx.y, q = y >> 1, s = y & 1;
    const uint32_t na = (uint32_t)(axis == 0 ? p.Ny : p.Nx), nb = (uint32_t)(axis == 2 ? p.Ny : p.Nz);
    const uint32_t plane = na * nb;
    const int n = axis == 0 ? p.Nx : (axis == 1 ? p.Ny : p.Nz);
    // load: second / second-to-last plane; save: first / last plane
    const int pos = save ? (s == 0 ? 0 : n - 1) : (s == 0 ? 1 : n - 2);
    double* const f = ff.f[q];
    const double* const g = save ? nullptr : ff.inner[q];
    double* const side = buf + (int64_t)y * plane;
    const uint32_t t = blockIdx.x * (uint32_t)kBlock + threadIdx.x;
    if (axis == 2 && p.sy == p.Nx) {
        double* const fp = f + (int64_t)pos * p.sz;
        if (vec) {
            const uint32_t c = 2 * t;
            if (c >= plane) return;
            double2* dst = reinterpret_cast<double2*>(save ? fp + c : side + c);
            const double2* src = reinterpret_cast<const double2*>(save ? side + c : fp + c);
            double2 v = *src;
            if (g) {
                const double2 w = *reinterpret_cast<const double2*>(g + (int64_t)pos * p.sz + c);
                v.x = interior3(p, (int)(c % na), (int)(c / na), pos) ? w.x : v.x;
                v.y = interior3(p, (int)((c + 1) % na), (int)((c + 1) / na), pos) ? w.y : v.y;
            }
            *dst = v;
        } else if (t < plane) {
            if (save)
                fp[t] = side[t];
            else
                side[t] = pick_inner(f, g, (int64_t)pos * p.sz + t, interior3(p, (int)(t % na), (int)(t / na), pos));
        }
        return;
    }
    if (t >= plane) return;
    const uint32_t a = t % na, b = t / na;
    int i, j, k;
    if (axis == 0) { i = pos; j = (int)a; k = (int)b; }
    else if (axis == 1) { i = (int)a; j = pos; k = (int)b; }
    else { i = (int)a; j = (int)b; k = pos; }
    const int64_t l = (int64_t)i + (int64_t)j * p.sy + (int64_t)k * p.sz;
    if (save)
        f[l] = side[t];
    else
        side[t] = pick_inner(f, g, l, interior3(p, i, j, k));
}

// Non-finite sentinel: the smallest cell index at which one of up to four
// state fields holds a NaN or an infinity, by atomicMin into *first (set to
// ~0 beforehand by k_fill_u64, a kernel rather than a memset so that the
// pair may be captured in a graph).  One atomic per wavefront that sees a
// bad cell.
__global__ void __launch_bounds__(kBlock) k_fill_u64(unsigned long long* dst, unsigned long long v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}
__global__ void __launch_bounds__(kBlock) k_nonfinite(DevParams p, const double* f0, const double* f1,
                                                      const double* f2, const double* f3,
                                                      unsigned long long* first) {
    const int64_t l = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool bad = false;
    if (l < p.n) {
        bad = !isfinite(f0[l]) || !isfinite(f1[l]);
        if (f2) bad = bad || !isfinite(f2[l]);
        if (f3) bad = bad || !isfinite(f3[l]);
    }
    const unsigned long long mask = __ballot(bad);
    if (mask == 0ull) return;
    // the lowest bad lane of the wavefront holds its smallest index
    const int lane = (int)(threadIdx.x & 63u);
    if (lane == __ffsll((long long)mask) - 1) atomicMin(first, (unsigned long long)l);
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

}  // namespace

hipError_t launch_entry(Entry e, const DevParams& p, const DevFields& F, hipStream_t s) {
    dim3 g = grid_for(p.n), b(kBlock);
    switch (e) {
        case Entry::S: hipLaunchKernelGGL(k_S, g, b, 0, s, p, F); break;
        case Entry::P_Xi_ro: hipLaunchKernelGGL(k_P_Xi_ro, g, b, 0, s, p, F); break;
        case Entry::U: hipLaunchKernelGGL(k_u, g, b, 0, s, p, F); break;
        case Entry::RoS: hipLaunchKernelGGL(k_roS, g, b, 0, s, p, F); break;
        case Entry::HE: hipLaunchKernelGGL(k_HE, g, b, 0, s, p, F); break;
        case Entry::E: hipLaunchKernelGGL(k_E, g, b, 0, s, p, F); break;
        case Entry::Newton: hipLaunchKernelGGL(k_newton, g, b, 0, s, p, F); break;
        case Entry::Boundary:
            hipLaunchKernelGGL(k_boundary_faces, grid_for(fused_face_cells(p)), b, 0, s, p, F);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_fused_step(const DevParams& p, const StepIO& io, hipStream_t s) {
    hipLaunchKernelGGL(k_fused_step, grid_for(p.n), dim3(kBlock), 0, s, p, io);
    return hipGetLastError();
}

// Workgroups of a kernel resident at once on the device (cached per kernel;
// one device type per process).
template <auto Kernel>
int resident_groups(int threads) {
    static int cached = 0;
    if (cached) return cached;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, Kernel, threads, 0) != hipSuccess || per_cu < 1)
        return 512;
    cached = cus * per_cu;
    return cached;
}

int fused_kchunk(const DevParams& p, int kchunk, int ty, int resident) {
    if (kchunk >= 1) return kchunk;
    // automatic: the planes per tile that minimise (rounds of resident
    // workgroups) x (planes per tile + the tile's extra coefficient planes)
    const int64_t nx = p.Nx >= 3 ? p.Nx - 2 : 1, ny = p.Ny - 2, nz = p.Nz >= 3 ? p.Nz - 2 : 1;
    const int64_t txy = ((nx + 63) / 64) * ((ny + ty - 1) / ty);
    int best = 16;
    double best_cost = -1.;
    for (int kc = 2; kc <= 16; ++kc) {
        const int64_t groups = txy * ((nz + kc - 1) / kc);
        const int64_t rounds = (groups + resident - 1) / resident;
        const double cost = (double)rounds * (std::min<int64_t>(kc, nz) + 0.5);
        if (best_cost < 0. || cost < best_cost) {
            best_cost = cost;
            best = kc;
        }
    }
    return best;
}

// The halo pack after a step whose launches wrote every new cell (nothing
// without io.pack).
template <int MODEL, int CLOSURE>
__global__ void __launch_bounds__(256) k_fused_pack(DevParams p, StepIO io) {
    fused_pack<MODEL, CLOSURE>(p, io, (int64_t)blockIdx.x * 256 + threadIdx.x);
}
template <int MODEL, int CLOSURE>
hipError_t launch_pack(const DevParams& p, const StepIO& io, hipStream_t s) {
    if (!io.pack) return hipSuccess;
    hipLaunchKernelGGL((k_fused_pack<MODEL, CLOSURE>), grid_for(2LL * p.Nx * p.Ny), dim3(256), 0, s, p, io);
    return hipGetLastError();
}
template <int MODEL>
hipError_t launch_pack_model(const DevParams& p, const StepIO& io, hipStream_t s) {
    if (p.closure == 0) return launch_pack<MODEL, 0>(p, io, s);
    if (p.corey_n == 2.) return launch_pack<MODEL, 2>(p, io, s);
    return launch_pack<MODEL, 1>(p, io, s);
}

template <int MODEL, int CLOSURE, int TY, bool CARRY>
hipError_t launch_z(const DevParams& p, const StepIO& io, int kchunk, hipStream_t s, const KernelEvents* ev) {
    const int nx = p.Nx >= 3 ? p.Nx - 2 : 1, ny = p.Ny - 2, nz = p.Nz >= 3 ? p.Nz - 2 : 1;
    constexpr auto kernel = k_fused_z<MODEL, CLOSURE, 64, TY, CARRY>;
    kchunk = fused_kchunk(p, kchunk, TY, resident_groups<kernel>(64 * TY));
    dim3 grid((nx + 63) / 64, (ny + TY - 1) / TY, (nz + kchunk - 1) / kchunk);
    hipError_t e;
    if (ev && (e = hipEventRecord(ev->begin, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(kernel, grid, dim3(64, TY), 0, s, p, io, kchunk);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev && (e = hipEventRecord(ev->end, s)) != hipSuccess) return e;
    if constexpr (!kFusedFaces<MODEL>) {
        const int64_t nfaces = fused_face_cells(p);
        if (io.pack)
            hipLaunchKernelGGL((k_fused_faces_pack<MODEL, CLOSURE>), grid_for(nfaces + 2LL * p.Nx * p.Ny), dim3(256),
                               0, s, p, io, nfaces);
        else
            hipLaunchKernelGGL((k_fused_faces<MODEL, CLOSURE>), grid_for(nfaces), dim3(256), 0, s, p, io);
        return hipGetLastError();
    }
    return launch_pack<MODEL, CLOSURE>(p, io, s);
}

template <int MODEL, int CLOSURE>
hipError_t launch_z_ty(const DevParams& p, const StepIO& io, const Tuning& t, hipStream_t s, const KernelEvents* ev) {
    const bool carry = t.fused_kernel == 1;
    // tile_y 0: 64 x 4 tiles for the two-phase Corey-2 step, 64 x 8 otherwise
    const int ty = t.tile_y ? t.tile_y : (MODEL == 2 && CLOSURE == 2 ? 4 : 8);
    if (ty == 8)
        return carry ? launch_z<MODEL, CLOSURE, 8, true>(p, io, t.kchunk, s, ev)
                     : launch_z<MODEL, CLOSURE, 8, false>(p, io, t.kchunk, s, ev);
    return carry ? launch_z<MODEL, CLOSURE, 4, true>(p, io, t.kchunk, s, ev)
                 : launch_z<MODEL, CLOSURE, 4, false>(p, io, t.kchunk, s, ev);
}

// CLOSURE ids of the fused kernels: 0 reference, 1 Corey with a run-time
// exponent, 2 Corey with exponent 2 (no pow in the kernel).
template <int MODEL>
hipError_t launch_z_model(const DevParams& p, const StepIO& io, const Tuning& t, hipStream_t s, const KernelEvents* ev) {
    if (p.closure == 0) return launch_z_ty<MODEL, 0>(p, io, t, s, ev);
    if (p.corey_n == 2.) return launch_z_ty<MODEL, 2>(p, io, t, s, ev);
    return launch_z_ty<MODEL, 1>(p, io, t, s, ev);
}

hipError_t launch_fused(const DevParams& p, const StepIO& io, const Tuning& t, hipStream_t s, const KernelEvents* ev) {
    if (t.fused_kernel == 0) {
        hipError_t e;
        if (ev && (e = hipEventRecord(ev->begin, s)) != hipSuccess) return e;
        if ((e = launch_fused_step(p, io, s)) != hipSuccess) return e;
        if (ev && (e = hipEventRecord(ev->end, s)) != hipSuccess) return e;
        return p.model == 3 ? launch_pack_model<3>(p, io, s) : launch_pack_model<2>(p, io, s);
    }
    return p.model == 3 ? launch_z_model<3>(p, io, t, s, ev) : launch_z_model<2>(p, io, t, s, ev);
}

hipError_t launch_fused_record(const DevParams& p, const StepIO& io, int kchunk, double* div, int64_t stride,
                               hipStream_t s) {
    const int nx = p.Nx >= 3 ? p.Nx - 2 : 1, ny = p.Ny - 2, nz = p.Nz >= 3 ? p.Nz - 2 : 1;
    if (p.model == 2) {
        // the Corey closures; the tiles write the boundary cells
        if (p.closure == 0) return hipErrorInvalidValue;
        kchunk = fused_kchunk(p, kchunk, 8, resident_groups<k_fused_z_record<2, 2, 64, 8, true>>(512));
        dim3 grid2((nx + 63) / 64, (ny + 7) / 8, (nz + kchunk - 1) / kchunk);
        if (p.corey_n == 2.)
            hipLaunchKernelGGL((k_fused_z_record<2, 2, 64, 8, true>), grid2, dim3(64, 8), 0, s, p, io, kchunk, div,
                               stride);
        else
            hipLaunchKernelGGL((k_fused_z_record<2, 1, 64, 8, true>), grid2, dim3(64, 8), 0, s, p, io, kchunk, div,
                               stride);
        return hipGetLastError();
    }
    if (p.model != 3) return hipErrorInvalidValue;
    kchunk = fused_kchunk(p, kchunk, 8, resident_groups<k_fused_z_record<3, 2, 64, 8, true>>(512));
    dim3 grid((nx + 63) / 64, (ny + 7) / 8, (nz + kchunk - 1) / kchunk);
    if (p.closure == 0) {
        hipLaunchKernelGGL((k_fused_z_record<3, 0, 64, 8, true>), grid, dim3(64, 8), 0, s, p, io, kchunk, div, stride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fused_faces<3, 0>), grid_for(fused_face_cells(p)), dim3(256), 0, s, p, io);
    } else if (p.corey_n == 2.) {
        hipLaunchKernelGGL((k_fused_z_record<3, 2, 64, 8, true>), grid, dim3(64, 8), 0, s, p, io, kchunk, div, stride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fused_faces<3, 2>), grid_for(fused_face_cells(p)), dim3(256), 0, s, p, io);
    } else {
        hipLaunchKernelGGL((k_fused_z_record<3, 1, 64, 8, true>), grid, dim3(64, 8), 0, s, p, io, kchunk, div, stride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_fused_faces<3, 1>), grid_for(fused_face_cells(p)), dim3(256), 0, s, p, io);
    }
    return hipGetLastError();
}

// roS (and, with U, u) of the interior cells on the z tiles, from the P, Xi
// and ro arrays of this step (three phases).
template <bool U>
hipError_t launch_roS_tiles(const DevParams& p, const DevFields& F, const double* dtm, hipStream_t s) {
    StepIO io{};
    io.Pw = F.f[F_P_w];
    io.Sw = F.f[F_S_w];
    io.Sn = F.f[F_S_n];
    io.Pn = F.f[F_P_n];
    io.Pg = F.f[F_P_g];
    io.m = F.f[F_m];
    io.K = F.f[F_K];
    io.dtm = dtm;
    const int xi[3] = {F_Xi_w, F_Xi_n, F_Xi_g}, ro[3] = {F_ro_w, F_ro_n, F_ro_g};
    const int rs[3] = {F_roS_w, F_roS_n, F_roS_g}, ro_old[3] = {F_roS_w_old, F_roS_n_old, F_roS_g_old};
    const int uf[3][3] = {{F_ux_w, F_uy_w, F_uz_w}, {F_ux_n, F_uy_n, F_uz_n}, {F_ux_g, F_uy_g, F_uz_g}};
    RoSStore<U> st{};
    st.dtm = dtm;
    for (int a = 0; a < 3; ++a) {
        io.Xi[a] = F.f[xi[a]];
        io.ro[a] = F.f[ro[a]];
        st.roS[a] = F.f[rs[a]];
        st.roS_old[a] = F.f[ro_old[a]];
        for (int x = 0; x < 3; ++x) st.u[a][x] = U ? F.f[uf[a][x]] : nullptr;
    }
    const int nx = p.Nx >= 3 ? p.Nx - 2 : 1, ny = p.Ny - 2, nz = p.Nz >= 3 ? p.Nz - 2 : 1;
    const int kchunk = fused_kchunk(p, 0, 8, resident_groups<k_roS_z<64, 8, true, U>>(512));
    dim3 grid((nx + 63) / 64, (ny + 7) / 8, (nz + kchunk - 1) / kchunk);
    hipLaunchKernelGGL((k_roS_z<64, 8, true, U>), grid, dim3(64, 8), 0, s, p, io, kchunk, st);
    return hipGetLastError();
}

// E_new of the interior cells on the z tiles; lam: the step's stored
// lambda_eff, or NULL to evaluate it per staged cell (the entry point).
hipError_t launch_energy_e_tiles(const DevParams& p, const DevFields& F, const double* lam, hipStream_t s) {
    EnergyEIO ea{};
    ea.lam = lam;
    ea.T = F.f[F_T];
    ea.E = F.f[F_E];
    ea.E_new = F.f[F_E_new];
    ea.m = F.f[F_m];
    ea.Sw = F.f[F_S_w];
    ea.Sn = F.f[F_S_n];
    ea.Sg = F.f[F_S_g];
    const int ro[3] = {F_ro_w, F_ro_n, F_ro_g}, H[3] = {F_H_w, F_H_n, F_H_g};
    const int ux[3] = {F_ux_w, F_ux_n, F_ux_g}, uy[3] = {F_uy_w, F_uy_n, F_uy_g}, uz[3] = {F_uz_w, F_uz_n, F_uz_g};
    for (int a = 0; a < 3; ++a) {
        ea.ro[a] = F.f[ro[a]];
        ea.H[a] = F.f[H[a]];
        ea.ux[a] = F.f[ux[a]];
        ea.uy[a] = F.f[uy[a]];
        ea.uzz[a] = F.f[p.ecfg ? uz[a] : uy[a]];
    }
    ea.lam_cfg = p.ecfg ? p.lam_eff : 0.;
    const int nx = p.Nx >= 3 ? p.Nx - 2 : 1, ny = p.Ny - 2, nz = p.Nz >= 3 ? p.Nz - 2 : 1;
    const int kc = fused_kchunk(p, 0, 8, resident_groups<k_energy_e_z<64, 8>>(512));
    dim3 ge((nx + 63) / 64, (ny + 7) / 8, (nz + kc - 1) / kc);
    hipLaunchKernelGGL((k_energy_e_z<64, 8>), ge, dim3(64, 8), 0, s, p, ea, kc);
    return hipGetLastError();
}

hipError_t launch_newton_energy(const DevParams& p, const DevFields& F, unsigned char* redo, hipStream_t s) {
    if (!redo) return launch_entry(Entry::Newton, p, F, s);
    hipLaunchKernelGGL(k_newton_energy, grid_for(p.n), dim3(kBlock), 0, s, p, F, redo);
    hipLaunchKernelGGL(k_newton_energy_redo, grid_for((p.n + 15) / 16), dim3(kBlock), 0, s, p, F, redo);
    return hipGetLastError();
}

hipError_t launch_roS_stored(const DevParams& p, const DevFields& F, const double* dtm, hipStream_t s) {
    if (p.model != 3) return hipErrorInvalidValue;
    return launch_roS_tiles<false>(p, F, dtm, s);
}

hipError_t launch_energy_step(const DevParams& p, const DevFields& F, double* lam, const double* dtm,
                              unsigned char* redo, hipStream_t s, const KernelEvents* ev) {
    hipError_t e;
    const dim3 g = grid_for(p.n), b(kBlock);
    if (ev && (e = hipEventRecord(ev->begin, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_energy_a, g, b, 0, s, p, F, lam);
    // u of the boundary cells here, of the interior cells in the roS tiles
    hipLaunchKernelGGL(k_energy_u_faces, grid_for(fused_face_cells(p)), b, 0, s, p, F);
    if ((e = launch_roS_tiles<true>(p, F, dtm, s)) != hipSuccess) return e;
    if ((e = launch_energy_e_tiles(p, F, lam, s)) != hipSuccess) return e;
    if ((e = launch_newton_energy(p, F, redo, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_boundary_faces, grid_for(fused_face_cells(p)), b, 0, s, p, F);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return ev ? hipEventRecord(ev->end, s) : hipSuccess;
}

hipError_t launch_dtm(const DevParams& p, const double* m, double* dtm, hipStream_t s) {
    hipLaunchKernelGGL(k_dtm, grid_for(p.n), dim3(kBlock), 0, s, p, m, dtm);
    return hipGetLastError();
}

hipError_t launch_copy_interior(const DevParams& p, const double* src, double* dst, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_interior, grid_for(p.n), dim3(kBlock), 0, s, p, src, dst);
    return hipGetLastError();
}

hipError_t launch_m_zero(const DevParams& p, const double* m, int* flag, hipStream_t s) {
    hipLaunchKernelGGL(k_m_zero, grid_for(p.n), dim3(kBlock), 0, s, p, m, flag);
    return hipGetLastError();
}

hipError_t launch_find_nonfinite(const DevParams& p, const double* const* f, unsigned long long* first,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(kBlock), 0, s, first, ~0ull);
    hipLaunchKernelGGL(k_nonfinite, grid_for(p.n), dim3(kBlock), 0, s, p, f[0], f[1], f[2], f[3], first);
    return hipGetLastError();
}

hipError_t launch_faces(const DevParams& p, double* const* fields, const double* const* inner, int nfields,
                        double* buf, int axis, int save, hipStream_t s) {
    if (nfields < 1 || nfields > kMaxFaceFields) return hipErrorInvalidValue;
    const int64_t na = axis == 0 ? p.Ny : p.Nx, nb = axis == 2 ? p.Ny : p.Nz, plane = na * nb;
    if (plane * 2 * nfields >= ((int64_t)1 << 31)) return hipErrorInvalidValue;
    FaceFields ff{};
    bool aligned = ((uintptr_t)buf & 15) == 0;
    for (int q = 0; q < nfields; ++q) {
        ff.f[q] = fields[q];
        ff.inner[q] = inner ? inner[q] : nullptr;
        aligned = aligned && ((uintptr_t)fields[q] & 15) == 0 && ((uintptr_t)ff.inner[q] & 15) == 0;
    }
    const int vec = axis == 2 && p.sy == p.Nx && plane % 2 == 0 && p.sz % 2 == 0 && aligned;
    const int64_t lanes = vec ? plane / 2 : plane;
    const dim3 grid((unsigned)((lanes + kBlock - 1) / kBlock), (unsigned)(2 * nfields));
    hipLaunchKernelGGL(k_faces, grid, dim3(kBlock), 0, s, p, ff, buf, axis, save, vec);
    return hipGetLastError();
}

hipError_t launch_face(const DevParams& p, double* field, const double* inner, double* buf, int axis, int pos,
                       int save, hipStream_t s) {
    int64_t na = axis == 0 ? p.Ny : p.Nx, nb = axis == 2 ? p.Ny : p.Nz;
    hipLaunchKernelGGL(k_face, grid_for(na * nb), dim3(kBlock), 0, s, p, field, inner, buf, axis, pos, save);
    return hipGetLastError();
}

}  // namespace capaz
