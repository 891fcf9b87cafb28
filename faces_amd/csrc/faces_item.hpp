// faces_item.hpp -- one self-contained work item of the periodic DSS step:
// S consecutive elements of one element row, loaded, summed in LDS and stored
// back by one block.  Used by the batched kernel, whose blocks take items of
// many grids (each with its own n and S) from a shared work list.
//
// The item's footprint is the one of the streaming kernel (faces_stream.hpp):
// OWN planes iz < n-1 of its elements, the y partner rows, the z partner
// plane, the yz partner row and the x seam column, laid out in LDS as
// StreamShape<N,S> describes.  Every location of a grid is a copy of exactly
// one node and each node is owned by exactly one item, so items of one grid
// (and of different grids) never write the same location.
//
// The last item of a row may hold fewer than S elements (mx need not be a
// multiple of S): slots of absent elements load and store nothing.
#pragma once

#include "faces_stream.hpp"

namespace faces {

// Per-grid constants of the batched step (device array, one per grid).
struct BatchGrid {
  double *u;
  long ys, ps;           // jy and jz strides (doubles)
  unsigned field_bytes;  // < kMaxBufferBytes
  int n, mx, my, mz;
  int S, nseg;           // elements per item (the row's last item may hold fewer), items per row
  int items;             // my*mz*nseg
  int first;             // index of the grid's first item in the work list
  double cost0, icost;   // estimated cost of the items before this grid; of one item
};

// Offsets of one item's footprint slots for the calling thread: load offset
// (kOOB for absent elements) and store offset (kOOB for positions the item
// must not store).  ecnt: elements present in the item.
template <int N, int S>
__device__ __forceinline__ int item_offsets(const BatchGrid &G, int it, unsigned (&ld)[StreamShape<N, S>::J],
                                            unsigned (&st)[StreamShape<N, S>::J])
{
  using SH = StreamShape<N, S>;
  constexpr int n = N, K = N - 1, n2 = N * N, JO = SH::JO;
  constexpr unsigned n3 = (unsigned)(N * N * N);
  constexpr unsigned kOOB = kOobBase;
  const int tid = threadIdx.x;
  const int sg = it % G.nseg;
  const int row = it / G.nseg;
  const int jy = row % G.my, jz = row / G.my;
  const int jym = (jy == 0) ? G.my - 1 : jy - 1;
  const int jzm = (jz == 0) ? G.mz - 1 : jz - 1;
  const int jx0 = sg * S;
  const int jxs = (jx0 == 0) ? G.mx - 1 : jx0 - 1;
  const int ecnt = min(S, G.mx - jx0);  // elements present in this item
  const unsigned r[4] = {(unsigned)((jz * G.ps + jy * G.ys) * 8), (unsigned)((jz * G.ps + jym * G.ys) * 8),
                         (unsigned)((jzm * G.ps + jy * G.ys) * 8), (unsigned)((jzm * G.ps + jym * G.ys) * 8)};
#pragma unroll
  for (int j = 0; j < JO; j++) {
    const int q = j * kStreamThreads + tid;
    const unsigned b = r[0] + jx0 * n3 * 8;
    ld[j] = b;
    st[j] = kOOB;
    if (q < SH::cOWN) {
      const int e = q / (K * n2), rr = q - e * (K * n2);
      const int ix = rr % n, iy = (rr / n) % n;
      ld[j] = e < ecnt ? b + (e * n3 + rr) * 8 : kOOB;
      if (e < ecnt && iy < K && !(e == ecnt - 1 && ix == K)) st[j] = ld[j];
    }
  }
#pragma unroll
  for (int jp = 0; jp < SH::JP; jp++) {
    const int j = JO + jp;
    int q = jp * kStreamThreads + tid;
    unsigned off = r[0] + jx0 * n3 * 8;
    bool keep = false;
    if (q < SH::cY) {
      const int e = q / (K * n), rr = q - e * (K * n);
      const int iz = rr / n, ix = rr - (rr / n) * n;
      off = e < ecnt ? r[1] + ((jx0 + e) * n3 + iz * n2 + K * n + ix) * 8 : kOOB;
      keep = e < ecnt && !(e == ecnt - 1 && ix == K);
    } else if ((q -= SH::cY) < SH::cZ) {
      const int e = q / (K * n), rr = q - e * (K * n);
      off = e < ecnt ? r[2] + ((jx0 + e) * n3 + K * n2 + rr) * 8 : kOOB;
      keep = e < ecnt && !(e == ecnt - 1 && rr % n == K);
    } else if ((q -= SH::cZ) < SH::cW) {
      const int e = q / n, ix = q - (q / n) * n;
      off = e < ecnt ? r[3] + ((jx0 + e) * n3 + K * n2 + K * n + ix) * 8 : kOOB;
      keep = e < ecnt && !(e == ecnt - 1 && ix == K);
    } else if ((q -= SH::cW) < SH::cS) {
      keep = true;
      const unsigned sb = jxs * n3;
      if (q < K * K) off = r[0] + (sb + (q / K) * n2 + (q % K) * n + K) * 8;
      else if (q < K * K + K) off = r[1] + (sb + (q - K * K) * n2 + K * n + K) * 8;
      else if (q < K * K + 2 * K) off = r[2] + (sb + K * n2 + (q - K * K - K) * n + K) * 8;
      else off = r[3] + (sb + K * n2 + K * n + K) * 8;
    }
    ld[j] = off;
    st[j] = keep ? off : kOOB;
  }
  return ecnt;
}

// The shared nodes of an item's first ecnt elements, summed in LDS: x pairs,
// then y, then z.
template <int N, int S>
__device__ __forceinline__ void item_sums(double *lds, int ecnt)
{
  using SH = StreamShape<N, S>;
  constexpr int n = N, K = N - 1, n2 = N * N, NN = SH::NN;
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < NN; i++) {
    const int k = i * kStreamThreads + tid;
    if (k >= ecnt * SH::nshared) continue;
    const int e = k / SH::nshared;
    int rr = k - e * SH::nshared;
    int ix, iy, iz;
    if (rr < K * K) {
      iz = 0; iy = rr / K; ix = rr - (rr / K) * K;
    } else if (rr < K * K + K * (K - 1)) {
      rr -= K * K;
      iz = 1 + rr / K; iy = 0; ix = rr - (rr / K) * K;
    } else {
      rr -= K * K + K * (K - 1);
      constexpr int K1 = (K > 1) ? K - 1 : 1;
      iz = 1 + rr / K1; iy = 1 + rr - (rr / K1) * K1; ix = 0;
    }
    const bool cx = ix == 0, cy = iy == 0, cz = iz == 0;
    const int p0 = e * (K * n2) + iz * n2 + iy * n + ix;
    const int p2 = SH::rY + e * (K * n) + iz * n + ix;
    const int p4 = SH::rZ + e * (K * n) + iy * n + ix;
    const int p6 = SH::rW + e * n + ix;
    const int p1 = e > 0 ? (e - 1) * (K * n2) + iz * n2 + iy * n + K : SH::rS + iz * K + iy;
    const int p3 = e > 0 ? SH::rY + (e - 1) * (K * n) + iz * n + K : SH::rS + K * K + iz;
    const int p5 = e > 0 ? SH::rZ + (e - 1) * (K * n) + iy * n + K : SH::rS + K * K + K + iy;
    const int p7 = e > 0 ? SH::rW + (e - 1) * n + K : SH::rS + K * K + 2 * K;
    const double a0 = cx ? lds[p0] + lds[p1] : lds[p0];
    double q0 = a0;
    if (cy) q0 = a0 + (cx ? lds[p2] + lds[p3] : lds[p2]);
    double res = q0;
    if (cz) {
      const double b0 = cx ? lds[p4] + lds[p5] : lds[p4];
      double q1 = b0;
      if (cy) q1 = b0 + (cx ? lds[p6] + lds[p7] : lds[p6]);
      res = q0 + q1;
    }
    lds[p0] = res;
    if (cx) lds[p1] = res;
    if (cy) {
      lds[p2] = res;
      if (cx) lds[p3] = res;
    }
    if (cz) {
      lds[p4] = res;
      if (cx) lds[p5] = res;
      if (cy) {
        lds[p6] = res;
        if (cx) lds[p7] = res;
      }
    }
  }
}

// A run of `count` items of one grid, item_of(0), item_of(1), ...: item t is
// summed while the loads of item t+1 are in flight.  Operations are issued
// per item as J loads (one item ahead), then J stores; J dropped stores
// before the loop enter it in that same state, so the waits are exact.
template <int N, int S, typename Order>
__device__ __forceinline__ void stream_run(const BatchGrid &G, const Order &item_of, int count, double *lds)
{
  using SH = StreamShape<N, S>;
  constexpr int J = SH::J;
  constexpr unsigned kOOB = kOobBase;
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const int tid = threadIdx.x;
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(G.u, 0, (int)G.field_bytes, 0x00020000);
  double v[J];
  {
    unsigned ld[J], st[J];
    item_offsets<N, S>(G, item_of(0), ld, st);
#pragma unroll
    for (int j = 0; j < J; j++)
      v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, ld[j], 0, 0));
  }
  static_assert(J <= kOobSlots, "dropped stores below the reserved offsets");
#pragma unroll
  for (int j = 0; j < J; j++)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, 0.0), rsrc, kOOB - 8 * j, 0, 0);
  int it = item_of(0);
  for (int t = 0; t < count; t++) {
    __syncthreads();  // the previous item's stores (or the previous run) have read LDS
#pragma unroll
    for (int j = 0; j < J; j++) lds[j * kStreamThreads + tid] = v[j];
    // the next item's loads (the last item re-reads its own locations: in
    // bounds, never used); offsets are recomputed where they are used, which
    // keeps only the loaded values live across the item
    const int nit = (t + 1 < count) ? item_of(t + 1) : it;
    {
      unsigned ld[J], st[J];
      item_offsets<N, S>(G, nit, ld, st);
#pragma unroll
      for (int j = 0; j < J; j++)
        v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, ld[j], 0, 0));
    }
    __syncthreads();
    item_sums<N, S>(lds, min(S, G.mx - (it % G.nseg) * S));
    __syncthreads();
    {
      unsigned ld[J], st[J];
      item_offsets<N, S>(G, it, ld, st);
#pragma unroll
      for (int j = 0; j < J; j++)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, lds[j * kStreamThreads + tid]),
                                              rsrc, st[j], 0, 0);
    }
    it = nit;
  }
  __syncthreads();  // LDS is reused by the block's next run
}

// One item, loaded, summed and stored by the calling block with nothing
// else in flight (the batched step's one-item-per-block form): the slots that
// hold no footprint location load nothing.
template <int N, int S>
__device__ __forceinline__ void item_once(const BatchGrid &G, int it, double *lds)
{
  using SH = StreamShape<N, S>;
  constexpr int J = SH::J;
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  const int tid = threadIdx.x;
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(G.u, 0, (int)G.field_bytes, 0x00020000);
  unsigned ld[J], st[J];
  const int ecnt = item_offsets<N, S>(G, it, ld, st);
  double v[J];
#pragma unroll
  for (int j = 0; j < J; j++) {
    const int q = j * kStreamThreads + tid;
    const bool used = (j < SH::JO) ? q < SH::cOWN : q - SH::JO * kStreamThreads < SH::cPART;
    v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, used ? ld[j] : kOobBase, 0, 0));
  }
#pragma unroll
  for (int j = 0; j < J; j++) lds[j * kStreamThreads + tid] = v[j];
  __syncthreads();
  item_sums<N, S>(lds, ecnt);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < J; j++)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, lds[j * kStreamThreads + tid]), rsrc,
                                          st[j], 0, 0);
}

}  // namespace faces
