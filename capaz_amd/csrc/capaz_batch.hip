// Batch kernels for gfx950: the fused step of many realizations in one
// launch.  A tile is the work unit of the single-domain z-marching kernel
// (TX x TY columns by kchunk planes, capaz_fused.h); the tiles of all
// realizations form one list, largest realization first, so that a
// workgroup that finishes early picks up the small realizations' tiles.
//
// The persistent form launches as many workgroups as fit on the GPU at once;
// each takes the next tile from a per-step counter (one atomic add per tile,
// broadcast through LDS) until the list is exhausted, so every workgroup
// exits once the counter passes the last tile.  The table of realizations
// is read-only and indexed by a wave-uniform value, so its fields load
// through scalar registers.
#include <hip/hip_runtime.h>

#include "capaz_batch.h"
#include "capaz_fused.h"

namespace capaz {

namespace {

// Occupancy bounds: the two-phase Corey-2 tiles of 256 threads (the 32 x 8
// and 64 x 4 shapes) fit in 94 VGPRs without a spill, 5 waves per SIMD
// against 4 at the compiler's own 106-108: the 64-realization ensemble went
// from 46.4 to 50.2 Gcell-steps/s; bound to 6 waves (80 VGPRs, 17 spilled)
// it ran at 44.5.  The three-phase Corey-2 tiles fit in 128 VGPRs without a
// spill against the compiler's own 129-131, 4 waves per SIMD against 3 (one
// 512-thread workgroup per CU for 64 x 8): a 24-realization three-phase
// batch (edges 32-96) 0.331 -> 0.305 ms per step on 32 x 8 tiles, 0.524 ->
// 0.376 ms on 64 x 8.
template <int MODEL, int CLOSURE, int TX, int TY>
constexpr int kBatchWaves =
    MODEL == 2 && CLOSURE == 2 && TX * TY == 256 ? 5 : (MODEL == 3 && CLOSURE == 2 ? 4 : 1);

template <int MODEL, int CLOSURE, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(kBatchWaves<MODEL, CLOSURE, TX, TY>)))
    k_batch_z(const BatchDomain* __restrict__ doms, const int* __restrict__ owner, int ntiles, int parity,
              int kchunk, int* counter) {
    __shared__ FusedLds<3 * MODEL, TX, TY> lds;
    __shared__ int claimed;
    // one workgroup per tile: each XCD (workgroups go to the 8 XCDs round
    // robin) takes a contiguous run of the tile list, so a realization's
    // neighbouring tiles share its L2
    int t = (int)blockIdx.x;
    if (!counter && gridDim.x % 8u == 0u) t = (int)((blockIdx.x % 8u) * (gridDim.x / 8u) + blockIdx.x / 8u);
    for (;;) {
        if (counter) {
            if (threadIdx.x == 0 && threadIdx.y == 0) claimed = atomicAdd(counter, 1);
            __syncthreads();
            t = __builtin_amdgcn_readfirstlane(claimed);
            __syncthreads();  // all have read `claimed` before it is overwritten
        }
        if (t < 0 || t >= ntiles) return;
        const BatchDomain* d = doms + owner[t];
        const int lt = t - d->tile_begin;
        const int bx = lt % d->tiles_x, r = lt / d->tiles_x;
        fused_z_tile<MODEL, CLOSURE, TX, TY, true>(d->p, d->io[parity], kchunk, bx, r % d->tiles_y, r / d->tiles_y,
                                                   lds, FusedStore<MODEL, kFusedFaces<MODEL>, CLOSURE>{d->io[parity]});
        if (!counter) return;
    }
}

template <int MODEL, int CLOSURE>
__global__ void __launch_bounds__(256)
    k_batch_faces(const BatchDomain* __restrict__ doms, const int* __restrict__ owner, int parity) {
    const BatchDomain* d = doms + owner[blockIdx.x];
    fused_face<MODEL, CLOSURE>(d->p, d->io[parity], (int64_t)((int)blockIdx.x - d->face_block_begin) * 256 + threadIdx.x);
}

__global__ void k_zero_ints(int* p, int n) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0;
}

template <int MODEL, int CLOSURE, int TX, int TY>
hipError_t step_shape(const BatchShape& sh, const BatchDomain* doms, const int* tile_owner, int ntiles,
                      const int* face_owner, int nface_blocks, int parity, int* counter, int grid, hipStream_t s) {
    if (ntiles > 0) {
        // (the one-per-tile grid is rounded up to a multiple of the 8 XCDs;
        // the extra workgroups find no tile and exit)
        const int g = sh.persistent ? grid : (ntiles + 7) / 8 * 8;
        hipLaunchKernelGGL((k_batch_z<MODEL, CLOSURE, TX, TY>), dim3(g), dim3(TX, TY), 0, s, doms, tile_owner, ntiles,
                           parity, sh.kchunk, sh.persistent ? counter : nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // two phases: the tiles write the boundary cells themselves
    if (nface_blocks > 0 && !kFusedFaces<MODEL>)
        hipLaunchKernelGGL((k_batch_faces<MODEL, CLOSURE>), dim3(nface_blocks), dim3(256), 0, s, doms, face_owner,
                           parity);
    return hipGetLastError();
}

template <int MODEL, int CLOSURE>
hipError_t step_closure(const BatchShape& sh, const BatchDomain* doms, const int* tile_owner, int ntiles,
                        const int* face_owner, int nface_blocks, int parity, int* counter, int grid, hipStream_t s) {
    if (sh.tx == 64 && sh.ty == 8)
        return step_shape<MODEL, CLOSURE, 64, 8>(sh, doms, tile_owner, ntiles, face_owner, nface_blocks, parity,
                                                 counter, grid, s);
    if (sh.tx == 64 && sh.ty == 4)
        return step_shape<MODEL, CLOSURE, 64, 4>(sh, doms, tile_owner, ntiles, face_owner, nface_blocks, parity,
                                                 counter, grid, s);
    return step_shape<MODEL, CLOSURE, 32, 8>(sh, doms, tile_owner, ntiles, face_owner, nface_blocks, parity, counter,
                                             grid, s);
}

template <int MODEL, int CLOSURE>
hipError_t occupancy_closure(const BatchShape& sh, int* n) {
    if (sh.tx == 64 && sh.ty == 8)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_batch_z<MODEL, CLOSURE, 64, 8>, 512, 0);
    if (sh.tx == 64 && sh.ty == 4)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_batch_z<MODEL, CLOSURE, 64, 4>, 256, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, k_batch_z<MODEL, CLOSURE, 32, 8>, 256, 0);
}

}  // namespace

int64_t batch_face_cells(const DevParams& p) { return fused_face_cells(p); }

bool batch_shape_supported(int tx, int ty) { return (tx == 64 && (ty == 8 || ty == 4)) || (tx == 32 && ty == 8); }

void batch_tiles(const DevParams& p, const BatchShape& sh, int* nx, int* ny, int* nz) {
    const int ix = p.Nx >= 3 ? p.Nx - 2 : 1, iy = p.Ny - 2, iz = p.Nz >= 3 ? p.Nz - 2 : 1;
    *nx = (ix + sh.tx - 1) / sh.tx;
    *ny = (iy + sh.ty - 1) / sh.ty;
    *nz = (iz + sh.kchunk - 1) / sh.kchunk;
}

hipError_t batch_occupancy(int model, int closure, const BatchShape& sh, int* n) {
    if (model == 3) {
        if (closure == 0) return occupancy_closure<3, 0>(sh, n);
        if (closure == 1) return occupancy_closure<3, 1>(sh, n);
        return occupancy_closure<3, 2>(sh, n);
    }
    if (closure == 0) return occupancy_closure<2, 0>(sh, n);
    if (closure == 1) return occupancy_closure<2, 1>(sh, n);
    return occupancy_closure<2, 2>(sh, n);
}

hipError_t launch_batch_step(int model, int closure, const BatchShape& sh, const BatchDomain* doms,
                             const int* tile_owner, int ntiles, const int* face_owner, int nface_blocks, int parity,
                             int* counter, int grid, hipStream_t s) {
#define CAPAZ_BATCH_ARGS sh, doms, tile_owner, ntiles, face_owner, nface_blocks, parity, counter, grid, s
    if (model == 3) {
        if (closure == 0) return step_closure<3, 0>(CAPAZ_BATCH_ARGS);
        if (closure == 1) return step_closure<3, 1>(CAPAZ_BATCH_ARGS);
        return step_closure<3, 2>(CAPAZ_BATCH_ARGS);
    }
    if (closure == 0) return step_closure<2, 0>(CAPAZ_BATCH_ARGS);
    if (closure == 1) return step_closure<2, 1>(CAPAZ_BATCH_ARGS);
    return step_closure<2, 2>(CAPAZ_BATCH_ARGS);
#undef CAPAZ_BATCH_ARGS
}

hipError_t launch_zero_ints(int* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_ints, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace capaz
