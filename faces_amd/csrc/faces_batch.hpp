// faces_batch.hpp -- device descriptors and launch helpers shared by the
// batched step (faces_batch.hip) and the serving loop (faces_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "faces_item.hpp"

namespace faces {

// One batched launch: grids sorted by first item; items [xr[x], xr[x+1]) go
// to the blocks that the hardware dispatches to XCD x (blocks x, x+8, ...).
constexpr int kXcds = 8;
struct BatchLaunch {
  const BatchGrid *grids;
  int ngrids;
  int nunits;
  int xr[kXcds + 1];
};

// Elements per item for a grid (largest S <= cap dividing mx with a kernel
// specialisation), 0 if none.
int batch_seg(int n, int mx, int cap);
int batch_lds_bytes(int n, int S);

// Fills S/nseg/items, sorts the grids by item cost, assigns first-item
// offsets and splits the work list into kXcds ranges of equal estimated
// cost; returns the LDS bytes and item count of the launch.
int plan_grids(std::vector<BatchGrid> &g, int cap, int *lds_bytes, int *nunits, int *xr);

// Persistent blocks for a launch: blocks_per_cu per CU, a multiple of kXcds.
int batch_blocks(int nunits, int blocks_per_cu);

// Enqueues the batched kernel (nblocks persistent blocks, a multiple of kXcds).
int launch_batch(const BatchLaunch &D, int nblocks, int lds_bytes, hipStream_t st);

}  // namespace faces
