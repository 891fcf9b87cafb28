// faces_batch.hpp -- device descriptors and launch helpers shared by the
// batched step (faces_batch.hip) and the serving loop (faces_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "faces_item.hpp"

namespace faces {

// One batched launch: grids sorted by first item.  The hardware dispatches
// block b to XCD b % 8; each XCD takes a contiguous share of the list of
// equal estimated cost (xs), as one item per block or, in the chunked form,
// as chunks of equal cost.
constexpr int kXcds = 8;
struct BatchLaunch {
  const BatchGrid *grids;  // null: the single grid g0 (passed by value)
  const int *firsts;       // grids[i].first as a compact array (follows the grids in their buffer)
  int ngrids;
  int nunits;
  double total;            // estimated cost of the whole list
  int xs[kXcds + 1];       // first item of each XCD's share; xs[kXcds] = nunits
  BatchGrid g0;
};

// Device (and pinned staging) buffer of a launch's descriptors for up to
// `cap` grids: the BatchGrid array, then the grids' first items as ints.
inline size_t batch_desc_bytes(int cap) { return (sizeof(BatchGrid) + sizeof(int)) * (size_t)cap; }

// Writes the descriptors of g into a buffer of that layout (grids, then
// their first items right after the last grid); returns the bytes written.
size_t batch_pack_desc(const std::vector<BatchGrid> &g, void *buf);

// Elements per item for a grid (largest S <= cap dividing mx with a kernel
// specialisation), 0 if none.
int batch_seg(int n, int mx, int cap);
int batch_lds_bytes(int n, int S);

// Fills S/nseg/items, orders the grids (dealt by item cost into one bin per
// XCD, so every XCD's share holds the same mix of item sizes) and assigns
// first-item offsets and cost prefixes; returns the LDS bytes and fills D's
// item count, total estimated cost and XCD shares (the caller sets grids,
// firsts and ngrids).
int plan_grids(std::vector<BatchGrid> &g, int cap, int *lds_bytes, BatchLaunch *D);

// blocks_per_cu value that selects the one-item-per-block kernel.
constexpr int kBatchOneItem = -1;

// Blocks of a launch: blocks_per_cu per CU (a chunk of the list each), or,
// for kBatchOneItem, kXcds times the largest XCD share (one block per item;
// the blocks past an XCD's share exit at once); a multiple of kXcds.
int batch_blocks(const BatchLaunch &D, int blocks_per_cu);

// Enqueues the batched kernel: nblocks blocks (a multiple of kXcds) taking a
// chunk of the list each, or (one_item) one item each.
int launch_batch(const BatchLaunch &D, int nblocks, int lds_bytes, hipStream_t st, bool one_item = false);

// One periodic grid (n <= 12, field below 4 GB) through the batched kernel:
// element segments of up to seg_cap elements, the row's last one possibly
// shorter.  The descriptor travels as a kernel argument: nothing is
// allocated or copied, so the launch can be captured in a graph.
int launch_single(double *u, int n, int mx, int my, int mz, long plane_stride, int seg_cap,
                  int blocks_per_cu, hipStream_t st);

}  // namespace faces
