// Batched fused steps: many independent domains (realizations of one
// physics, each with its own grid and fields) advanced by one launch per
// kernel and step.  Private to the library.
#pragma once

#include <hip/hip_runtime.h>

#include "capaz_bodies.h"

namespace capaz {

// One realization as the batch kernels read it (device table, one entry per
// realization, in launch order: largest interior first).
struct BatchDomain {
    DevParams p;
    StepIO io[2];         // [0]: current arrays -> second arrays, [1]: back
    int tiles_x, tiles_y, tiles_z;
    int tile_begin;       // first tile of this realization in the tile order
    int face_block_begin; // first 256-cell block of its boundary cells
    int face_blocks;
};

// Tile shape of the batch kernels: TX x TY columns, kchunk planes per tile;
// persistent 1: a grid that fills the GPU once and takes tiles from a queue
// (one atomic counter per step), 0: one workgroup per tile.
struct BatchShape {
    int tx = 64, ty = 8, kchunk = 16, persistent = 1;
};

bool batch_shape_supported(int tx, int ty);
// Boundary cells of one realization (threads of its face pass).
int64_t batch_face_cells(const DevParams& p);
// Tiles of one realization for a shape.
void batch_tiles(const DevParams& p, const BatchShape& sh, int* nx, int* ny, int* nz);
// Workgroups of the interior kernel that fit one CU at once.
hipError_t batch_occupancy(int model, int closure, const BatchShape& sh, int* blocks_per_cu);
// One step of every realization: the interior tiles, then the boundary cells.
// counter: a zeroed int for the persistent kernel's queue (NULL otherwise);
// grid: workgroups of the persistent kernel.
hipError_t launch_batch_step(int model, int closure, const BatchShape& sh, const BatchDomain* doms,
                             const int* tile_owner, int ntiles, const int* face_owner, int nface_blocks, int parity,
                             int* counter, int grid, hipStream_t s);
// Zeroes n ints with a kernel (graph-safe: no memset node).
hipError_t launch_zero_ints(int* p, int n, hipStream_t s);

}  // namespace capaz
