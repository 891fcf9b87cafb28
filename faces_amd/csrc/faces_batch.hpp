// faces_batch.hpp -- device descriptors and launch helpers shared by the
// batched step (faces_batch.hip) and the serving loop (faces_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "faces_item.hpp"

namespace faces {

// One batched launch: grids sorted by first item, a work counter pair.
struct BatchLaunch {
  const BatchGrid *grids;
  int ngrids;
  int nunits;
  unsigned *counter;  // [0] next item, [1] finished blocks (zero between launches)
};

// Elements per item for a grid (largest S <= cap dividing mx with a kernel
// specialisation), 0 if none.
int batch_seg(int n, int mx, int cap);
int batch_lds_bytes(int n, int S);

// Fills S/nseg/items, sorts the grids by item cost and assigns first-item
// offsets; returns the LDS bytes and item count of the launch.
int plan_grids(std::vector<BatchGrid> &g, int cap, int *lds_bytes, int *nunits);

// Enqueues the batched kernel (nblocks persistent blocks).
int launch_batch(const BatchLaunch &D, int nblocks, int lds_bytes, hipStream_t st);

}  // namespace faces
