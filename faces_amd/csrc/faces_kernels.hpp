// faces_kernels.hpp -- CDNA4 (gfx950) HIP kernels for the Faces exchange +
// direct-stiffness-summation (DSS) step.
//
// Field layout (reference gpu/DArray.h:36-38, gpu/Array.h:24-26):
//   u(ix,iy,iz,jx,jy,jz) = u[ix + n*(iy + n*(iz + n*(jx + mx*jy))) + jz*ps]
// with ix fastest, n^3 FP64 points per element, and ps the jz-plane stride
// (n^3*mx*my when dense, larger for the reference's page-padded DArray,
// gpu/DArray.h:95-99).
//
// Summation semantics.  Every unique mesh node has 1, 2, 4 or 8 copies; the
// step replaces each copy by the sum of all copies, grouped x first, then y,
// then z:  ((c000+c100)+(c010+c110)) + ((c001+c101)+(c011+c111)).  That is the
// order the reference's host check evaluates (gpu/Mugs.cc:136-203: x-pass,
// y-pass, z-pass of pairwise sums); addition is commutative, so the grouping
// alone fixes the bits, and every kernel here reproduces it exactly.
//
// Ownership.  Each unique node is handled by exactly one lane, which reads all
// of the node's copies and writes all of them.  A memory location is a copy
// of exactly one node, so the in-place update is free of races without any
// staging buffer, and every byte of the field is read once and written once.
// Per axis the lane of a node at global index g (element j, local index i,
// g = j*(n-1) + i) sees:
//   i != 0            one copy (j, i)
//   i == 0, interior  two copies (j, 0) and (j-1, n-1)
//   periodic wrap     g == 0 pairs (0, 0) with (m-1, n-1)
//   rank surface      g == 0 or g == m*(n-1) in a non-periodic axis: one local
//                     copy plus, when the neighbour exists, one received copy
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace faces {

// ---------------------------------------------------------------------------
// descriptors passed by value to kernels
// ---------------------------------------------------------------------------

struct Geom {
  double *u;
  long ps;  // jz-plane stride (doubles)
  long ys;  // jy stride = mx*n^3
  int n, mx, my, mz;
};

// Axis boundary: wrap = periodic self-pairing inside the field; otherwise
// lo/hi say whether a received copy exists on that side.
struct Axis {
  int wrap, lo, hi;
};

struct Bound {
  Axis x, y, z;
};

// Offsets (doubles) of the 26 message slots inside one message arena.  Slot
// numbering follows the reference neighbour tables
// (gpu/v-26-messages/Faces.cc:37-75):
//   0..5   faces x-,x+,y-,y+,z-,z+
//   6..9   x-edges, e = sy + 2*sz          (iedge_[0..3])
//   10..13 y-edges, e = sx + 2*sz          (iedge_[4..7])
//   14..17 z-edges, e = sx + 2*sy          (iedge_[8..11])
//   18..25 corners, c = sx + 2*sy + 4*sz   (icorner_)
struct Slots {
  long off[27];
};

__host__ __device__ inline int opposite_slot(int k)
{
  if (k < 6) return (k & ~1) | (1 - (k & 1));
  if (k < 18) return 6 + 4 * ((k - 6) / 4) + (3 - (k - 6) % 4);
  return 18 + (7 - (k - 18));
}

// ---------------------------------------------------------------------------
// per-axis copy sets
// ---------------------------------------------------------------------------

struct AxCopies {
  int cnt;
  int j[2], i[2];
  int rem[2];  // copy lives in the received message of side s
  int s[2];
};

template <int N>
__device__ __forceinline__ int nval(int n)
{
  return N > 0 ? N : n;
}

template <int N>
__device__ __forceinline__ void axis_copies(int g, int m, int n_rt, Axis a, AxCopies &c)
{
  const int n = nval<N>(n_rt);
  const int K = n - 1;
  c.rem[0] = 0;
  c.rem[1] = 0;
  c.s[0] = 0;
  c.s[1] = 0;
  if (a.wrap) {
    const int j = g / K, i = g - (g / K) * K;
    c.j[0] = j;
    c.i[0] = i;
    if (i == 0) {
      c.cnt = 2;
      c.j[1] = (j == 0) ? m - 1 : j - 1;
      c.i[1] = K;
    } else {
      c.cnt = 1;
    }
    return;
  }
  const int G = m * K + 1;
  if (g == 0) {
    c.j[0] = 0;
    c.i[0] = 0;
    c.cnt = 1;
    if (a.lo) {
      c.cnt = 2;
      c.j[1] = 0;
      c.i[1] = 0;
      c.rem[1] = 1;
      c.s[1] = 0;
    }
  } else if (g == G - 1) {
    c.j[0] = m - 1;
    c.i[0] = K;
    c.cnt = 1;
    if (a.hi) {
      c.cnt = 2;
      c.j[1] = m - 1;
      c.i[1] = K;
      c.rem[1] = 1;
      c.s[1] = 1;
    }
  } else {
    const int j = g / K, i = g - (g / K) * K;
    c.j[0] = j;
    c.i[0] = i;
    if (i == 0) {
      c.cnt = 2;
      c.j[1] = j - 1;
      c.i[1] = K;
    } else {
      c.cnt = 1;
    }
  }
}

// Index of a received copy inside the message arena.  Face, edge and corner
// index maps are those of the reference's unpack
// (gpu/v-26-messages/Faces.cc:554-669): an x-face message is ordered
// (iy,iz,jy,jz), a y-face (ix,iz,jx,jz), a z-face (ix,iy,jx,jy); an x-edge
// (ix,jx), a y-edge (iy,jy), a z-edge (iz,jz).
__device__ __forceinline__ long recv_index(const Slots &sl, int n, int mx, int my, int rx, int ry,
                                           int rz, int sx, int sy, int sz, int ix, int iy, int iz,
                                           int jx, int jy, int jz)
{
  const long nn = (long)n * n;
  if (rx && !ry && !rz) return sl.off[0 + sx] + iy + n * iz + nn * (jy + (long)my * jz);
  if (!rx && ry && !rz) return sl.off[2 + sy] + ix + n * iz + nn * (jx + (long)mx * jz);
  if (!rx && !ry && rz) return sl.off[4 + sz] + ix + n * iy + nn * (jx + (long)mx * jy);
  if (!rx && ry && rz) return sl.off[6 + sy + 2 * sz] + ix + (long)n * jx;
  if (rx && !ry && rz) return sl.off[10 + sx + 2 * sz] + iy + (long)n * jy;
  if (rx && ry && !rz) return sl.off[14 + sx + 2 * sy] + iz + (long)n * jz;
  return sl.off[18 + sx + 2 * sy + 4 * sz];
}

// ---------------------------------------------------------------------------
// generic node: any boundary, received copies allowed (surface kernel)
// ---------------------------------------------------------------------------

template <int N>
__device__ __forceinline__ void dss_node_general(const Geom &G, const Bound &B,
                                                 const double *__restrict__ recv, const Slots &sl,
                                                 int gx, int gy, int gz)
{
  const int n = nval<N>(G.n);
  AxCopies X, Y, Z;
  axis_copies<N>(gx, G.mx, n, B.x, X);
  axis_copies<N>(gy, G.my, n, B.y, Y);
  axis_copies<N>(gz, G.mz, n, B.z, Z);
  const long n3 = (long)n * n * n;
  double v[2][2][2];
#pragma unroll
  for (int az = 0; az < 2; az++)
#pragma unroll
    for (int ay = 0; ay < 2; ay++)
#pragma unroll
      for (int ax = 0; ax < 2; ax++) {
        if (ax < X.cnt && ay < Y.cnt && az < Z.cnt) {
          const int rx = X.rem[ax], ry = Y.rem[ay], rz = Z.rem[az];
          if (!rx && !ry && !rz) {
            const long a = X.j[ax] * n3 + X.i[ax] + Y.j[ay] * G.ys + (long)Y.i[ay] * n +
                           Z.j[az] * G.ps + (long)Z.i[az] * n * n;
            v[az][ay][ax] = G.u[a];
          } else {
            v[az][ay][ax] = recv[recv_index(sl, n, G.mx, G.my, rx, ry, rz, X.s[ax], Y.s[ay], Z.s[az],
                                            X.i[ax], Y.i[ay], Z.i[az], X.j[ax], Y.j[ay], Z.j[az])];
          }
        }
      }
  double q[2];
#pragma unroll
  for (int az = 0; az < 2; az++) {
    double p0 = (X.cnt == 2) ? v[az][0][0] + v[az][0][1] : v[az][0][0];
    double p1 = (X.cnt == 2) ? v[az][1][0] + v[az][1][1] : v[az][1][0];
    q[az] = (Y.cnt == 2) ? p0 + p1 : p0;
  }
  const double r = (Z.cnt == 2) ? q[0] + q[1] : q[0];
#pragma unroll
  for (int az = 0; az < 2; az++)
#pragma unroll
    for (int ay = 0; ay < 2; ay++)
#pragma unroll
      for (int ax = 0; ax < 2; ax++) {
        if (ax < X.cnt && ay < Y.cnt && az < Z.cnt && !X.rem[ax] && !Y.rem[ay] && !Z.rem[az]) {
          const long a = X.j[ax] * n3 + X.i[ax] + Y.j[ay] * G.ys + (long)Y.i[ay] * n +
                         Z.j[az] * G.ps + (long)Z.i[az] * n * n;
          G.u[a] = r;
        }
      }
}

// ---------------------------------------------------------------------------
// main-box node: g < m*(n-1) on every axis, all copies local
// ---------------------------------------------------------------------------
// (jx,ix) etc. are the node's owning element and local index (i < n-1).  The
// low partner of i == 0 is element j-1, or m-1 for the periodic wrap (only
// reached when the axis wraps; non-periodic callers skip g == 0).

template <int N>
__device__ __forceinline__ void dss_node_box(const Geom &G, int jx, int ix, int jy, int iy, int jz,
                                             int iz)
{
  const int n = nval<N>(G.n);
  const long n3 = (long)n * n * n;
  const int cx = (ix == 0) ? 2 : 1, cy = (iy == 0) ? 2 : 1, cz = (iz == 0) ? 2 : 1;
  const int jxm = (jx == 0) ? G.mx - 1 : jx - 1;
  const int jym = (jy == 0) ? G.my - 1 : jy - 1;
  const int jzm = (jz == 0) ? G.mz - 1 : jz - 1;
  long xo[2], yo[2], zo[2];
  xo[0] = jx * n3 + ix;
  xo[1] = jxm * n3 + (n - 1);
  yo[0] = jy * G.ys + (long)iy * n;
  yo[1] = jym * G.ys + (long)(n - 1) * n;
  zo[0] = jz * G.ps + (long)iz * n * n;
  zo[1] = jzm * G.ps + (long)(n - 1) * n * n;
  double *__restrict__ u = G.u;
  double v[2][2][2];
#pragma unroll
  for (int az = 0; az < 2; az++)
#pragma unroll
    for (int ay = 0; ay < 2; ay++)
#pragma unroll
      for (int ax = 0; ax < 2; ax++)
        if (ax < cx && ay < cy && az < cz) v[az][ay][ax] = u[zo[az] + yo[ay] + xo[ax]];
  double q[2];
#pragma unroll
  for (int az = 0; az < 2; az++) {
    double p0 = (cx == 2) ? v[az][0][0] + v[az][0][1] : v[az][0][0];
    double p1 = (cx == 2) ? v[az][1][0] + v[az][1][1] : v[az][1][0];
    q[az] = (cy == 2) ? p0 + p1 : p0;
  }
  const double r = (cz == 2) ? q[0] + q[1] : q[0];
#pragma unroll
  for (int az = 0; az < 2; az++)
#pragma unroll
    for (int ay = 0; ay < 2; ay++)
#pragma unroll
      for (int ax = 0; ax < 2; ax++)
        if (ax < cx && ay < cy && az < cz) u[zo[az] + yo[ay] + xo[ax]] = r;
}

// ---------------------------------------------------------------------------
// row-blocked main-box kernel
// ---------------------------------------------------------------------------
// One work item = one element row segment (jy, jz, jx in [jx0, jx1)); its lanes
// own the (n-1)^3 lower nodes of each element in the segment, enumerated
// element-major so a wavefront's own copies are consecutive in memory.  The
// x-partner copies live in the previous element of the same row (cached in
// L1); the y/z partners live in neighbouring rows that the XCD-aware item
// order schedules on the same XCD at about the same time, so their sectors
// merge in that XCD's L2.

struct RowItems {
  int my, mz;      // rows = my*mz
  int segs;        // segments per row
  int seg_len;     // elements per segment
  int xcd_remap;   // 1: contiguous item ranges per XCD
  int nitems;      // my*mz*segs
};

__device__ __forceinline__ int item_of_block(const RowItems &R, int b, int nblocks)
{
  if (!R.xcd_remap) return b;
  // Work-groups are dispatched round-robin over the 8 XCDs (block b -> XCD
  // b%8).  Give XCD x the contiguous item range [x*q + min(x,rem), ...).
  const int x = b & 7, k = b >> 3;
  const int q = R.nitems >> 3, rem = R.nitems & 7;
  const int cnt = q + (x < rem ? 1 : 0);
  if (k >= cnt) return -1;
  (void)nblocks;
  return x * q + (x < rem ? x : rem) + k;
}

template <int N, bool SKIP_SURFACE>
__device__ __forceinline__ void box_segment(const Geom &G, int jy, int jz, int jx0, int jx1, int tid,
                                            int nthreads)
{
  const int n = nval<N>(G.n);
  const int K = n - 1;
  const int K2 = K * K, K3 = K2 * K;
  const int total = (jx1 - jx0) * K3;
  for (int t = tid; t < total; t += nthreads) {
    const int e = t / K3;
    const int r = t - e * K3;
    const int iz = r / K2;
    const int r2 = r - iz * K2;
    const int iy = r2 / K;
    const int ix = r2 - iy * K;
    const int jx = jx0 + e;
    if (SKIP_SURFACE && ((jx == 0 && ix == 0) || (jy == 0 && iy == 0) || (jz == 0 && iz == 0)))
      continue;
    dss_node_box<N>(G, jx, ix, jy, iy, jz, iz);
  }
}

template <int N, bool SKIP_SURFACE>
__global__ __launch_bounds__(1024) void k_box(Geom G, RowItems R)
{
  const int it = item_of_block(R, blockIdx.x, gridDim.x);
  if (it < 0 || it >= R.nitems) return;
  const int seg = it % R.segs;
  const int row = it / R.segs;
  const int jy = row % R.my, jz = row / R.my;
  const int jx0 = seg * R.seg_len;
  const int jx1 = min(G.mx, jx0 + R.seg_len);
  box_segment<N, SKIP_SURFACE>(G, jy, jz, jx0, jx1, threadIdx.x, blockDim.x);
}

// ---------------------------------------------------------------------------
// LDS-staged element kernel (periodic field, every axis wraps)
// ---------------------------------------------------------------------------
// A work item is a segment of S consecutive elements [jx0, jx0+S) of one
// element row (jy, jz), in memory order.  It owns the lower-corner nodes
// (ix, iy, iz < n-1) of its elements.  Their copies are:
//   OWN  the segment's elements, offsets [0, (n-1)*n^2) of each (planes
//        iz < n-1; the iy = n-1 rows in there are loaded but belong to the
//        next row's item and are not stored)
//   YP   row (jy-1, jz):   x-rows iy = n-1 of planes iz < n-1
//   ZP   row (jy, jz-1):   plane iz = n-1, x-rows iy < n-1
//   WP   row (jy-1, jz-1): plane iz = n-1, x-row iy = n-1
//   SEAM the ix = n-1 column of element jx0-1 (m-1 for jx0 = 0) in the four
//        rows above: the x partners of the segment's first element
// The positions ix = n-1 of the segment's last element belong to the next
// segment.  Loads and stores follow memory order, so the bulk (OWN, 86% of
// the field at n=7) streams as nearly contiguous 2352-byte runs; YP/ZP/WP
// are touched by the items of the neighbouring rows shortly before, and
// XCD-contiguous item order keeps them in that XCD's L2.

struct ElemItems {
  int my, mz;
  int seg;        // elements per item
  int nseg;       // items per row
  int nitems;     // my*mz*nseg
  int xcd_remap;
};

template <int N>
struct ElemLayout {
  static constexpr int n = N, K = N - 1, n2 = N * N, n3 = N * N * N;
  int S;           // elements in this item
  int o_y, o_z, o_w, o_s, total;
  __device__ __forceinline__ ElemLayout(int s)
  {
    S = s;
    o_y = S * K * n2;
    o_z = o_y + S * K * n;
    o_w = o_z + S * K * n;
    o_s = o_w + S * n;
    total = o_s + n2;
  }
};

// Footprint position t -> field index; `keep` = the item stores it.
template <int N>
__device__ __forceinline__ long elem_global(const ElemLayout<N> &L, int t, int jx0, int jxs, long own,
                                            long ym, long zm, long yzm, bool &keep)
{
  constexpr int n = N, K = N - 1, n2 = N * N;
  constexpr long n3 = (long)N * N * N;
  keep = true;
  if (t < L.o_y) {
    const int e = t / (K * n2), r = t - e * (K * n2);
    const int ix = r % n, iy = (r / n) % n;
    keep = (iy < K) && !(e == L.S - 1 && ix == K);
    return own + (jx0 + e) * n3 + r;
  }
  if (t < L.o_z) {
    t -= L.o_y;
    const int e = t / (K * n), r = t - e * (K * n);
    const int iz = r / n, ix = r - (r / n) * n;
    keep = !(e == L.S - 1 && ix == K);
    return ym + (jx0 + e) * n3 + iz * n2 + K * n + ix;
  }
  if (t < L.o_w) {
    t -= L.o_z;
    const int e = t / (K * n), r = t - e * (K * n);
    const int ix = r % n;
    keep = !(e == L.S - 1 && ix == K);
    return zm + (jx0 + e) * n3 + K * n2 + r;
  }
  if (t < L.o_s) {
    t -= L.o_w;
    const int e = t / n, ix = t - (t / n) * n;
    keep = !(e == L.S - 1 && ix == K);
    return yzm + (jx0 + e) * n3 + K * n2 + K * n + ix;
  }
  t -= L.o_s;
  const long sb = jxs * n3;
  if (t < K * K) {
    const int iz = t / K, iy = t - (t / K) * K;
    return own + sb + iz * n2 + iy * n + K;
  }
  t -= K * K;
  if (t < K) return ym + sb + t * n2 + K * n + K;
  t -= K;
  if (t < K) return zm + sb + K * n2 + t * n + K;
  return yzm + sb + K * n2 + K * n + K;
}

template <int N, int BATCH>
__global__ __launch_bounds__(512) void k_elem_lds(Geom G, ElemItems I)
{
  extern __shared__ double lds[];
  constexpr int n = N, K = N - 1, n2 = N * N;
  const int it = item_of_block(RowItems{I.my, I.mz, 1, 1, I.xcd_remap, I.nitems}, blockIdx.x,
                               gridDim.x);
  if (it < 0 || it >= I.nitems) return;
  const int sg = it % I.nseg;
  const int row = it / I.nseg;
  const int jy = row % G.my, jz = row / G.my;
  const int jym = (jy == 0) ? G.my - 1 : jy - 1;
  const int jzm = (jz == 0) ? G.mz - 1 : jz - 1;
  const int jx0 = sg * I.seg;
  const int S = min(I.seg, G.mx - jx0);
  const int jxs = (jx0 == 0) ? G.mx - 1 : jx0 - 1;
  double *__restrict__ u = G.u;
  const long own = jz * G.ps + jy * G.ys;
  const long ym = jz * G.ps + jym * G.ys;
  const long zm = jzm * G.ps + jy * G.ys;
  const long yzm = jzm * G.ps + jym * G.ys;
  const ElemLayout<N> L(S);
  const int tid = threadIdx.x, nt = blockDim.x;

  // ---- load the footprint into LDS, memory order
  for (int t0 = tid; t0 < L.total; t0 += nt * BATCH) {
    double v[BATCH];
#pragma unroll
    for (int b = 0; b < BATCH; b++) {
      const int t = t0 + b * nt;
      bool keep;
      if (t < L.total) v[b] = u[elem_global<N>(L, t, jx0, jxs, own, ym, zm, yzm, keep)];
    }
#pragma unroll
    for (int b = 0; b < BATCH; b++) {
      const int t = t0 + b * nt;
      if (t < L.total) lds[t] = v[b];
    }
  }
  __syncthreads();

  // ---- sum the owned nodes (e, iz, iy, ix < n-1) in LDS
  for (int t = tid; t < S * K * K * K; t += nt) {
    const int e = t / (K * K * K);
    const int r = t - e * (K * K * K);
    const int iz = r / (K * K);
    const int r2 = r - iz * (K * K);
    const int iy = r2 / K, ix = r2 - (r2 / K) * K;
    const int cx = (ix == 0), cy = (iy == 0), cz = (iz == 0);
    // LDS positions p[az][ay][ax]; ax = 1 is the previous element's
    // ix = n-1 copy (the seam column when e == 0).
    int p[2][2][2];
    p[0][0][0] = e * (K * n2) + iz * n2 + iy * n + ix;
    p[0][1][0] = L.o_y + e * (K * n) + iz * n + ix;
    p[1][0][0] = L.o_z + e * (K * n) + iy * n + ix;
    p[1][1][0] = L.o_w + e * n + ix;
    if (e > 0) {
      p[0][0][1] = (e - 1) * (K * n2) + iz * n2 + iy * n + K;
      p[0][1][1] = L.o_y + (e - 1) * (K * n) + iz * n + K;
      p[1][0][1] = L.o_z + (e - 1) * (K * n) + iy * n + K;
      p[1][1][1] = L.o_w + (e - 1) * n + K;
    } else {
      p[0][0][1] = L.o_s + iz * K + iy;
      p[0][1][1] = L.o_s + K * K + iz;
      p[1][0][1] = L.o_s + K * K + K + iy;
      p[1][1][1] = L.o_s + K * K + 2 * K;
    }
    // x pairs, then y, then z (absent copies are never read)
    const double a0 = cx ? lds[p[0][0][0]] + lds[p[0][0][1]] : lds[p[0][0][0]];
    double q0 = a0;
    if (cy) q0 = a0 + (cx ? lds[p[0][1][0]] + lds[p[0][1][1]] : lds[p[0][1][0]]);
    double res = q0;
    if (cz) {
      const double b0 = cx ? lds[p[1][0][0]] + lds[p[1][0][1]] : lds[p[1][0][0]];
      double q1 = b0;
      if (cy) q1 = b0 + (cx ? lds[p[1][1][0]] + lds[p[1][1][1]] : lds[p[1][1][0]]);
      res = q0 + q1;
    }
    lds[p[0][0][0]] = res;
    if (cx) lds[p[0][0][1]] = res;
    if (cy) {
      lds[p[0][1][0]] = res;
      if (cx) lds[p[0][1][1]] = res;
    }
    if (cz) {
      lds[p[1][0][0]] = res;
      if (cx) lds[p[1][0][1]] = res;
      if (cy) {
        lds[p[1][1][0]] = res;
        if (cx) lds[p[1][1][1]] = res;
      }
    }
  }
  __syncthreads();

  // ---- store the owned copies back, memory order
  for (int t0 = tid; t0 < L.total; t0 += nt * BATCH) {
#pragma unroll
    for (int b = 0; b < BATCH; b++) {
      const int t = t0 + b * nt;
      if (t < L.total) {
        bool keep;
        const long g = elem_global<N>(L, t, jx0, jxs, own, ym, zm, yzm, keep);
        if (keep) u[g] = lds[t];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// surface kernel: the unique nodes on the rank surface
// ---------------------------------------------------------------------------
// Surface values per axis: {0} when periodic, {0, G-1} otherwise.  Nodes are
// enumerated in three disjoint parts -- z-surface planes; y-surface planes off
// the z-surface; x-surface planes off both -- the same partition as the
// reference's three unpack kernels (gpu/v-26-messages/Faces.cc:554-669).

struct SurfDesc {
  int Gx, Gy, Gz;     // owned node counts per axis
  int nsx, nsy, nsz;  // surface values per axis (1 or 2)
  long cz, cy, cx;    // node counts of the three parts
};

__host__ __device__ inline int surf_value(int k, int G)
{
  return k == 0 ? 0 : G - 1;
}

// Surface node t of the enumeration above: its copies (and received copies)
// summed and written back.
template <int N>
__device__ __forceinline__ void surface_node(const Geom &G, const Bound &B, const double *__restrict__ recv,
                                             const Slots &sl, const SurfDesc &S, long t)
{
  int gx, gy, gz;
  if (t < S.cz) {
    const long pl = (long)S.Gx * S.Gy;
    const int k = (int)(t / pl);
    const long r = t - k * pl;
    gz = surf_value(k, S.Gz);
    gy = (int)(r / S.Gx);
    gx = (int)(r - (long)gy * S.Gx);
  } else if (t < S.cz + S.cy) {
    const long tt = t - S.cz;
    const int niz = S.Gz - S.nsz;
    const long pl = (long)S.Gx * niz;
    const int k = (int)(tt / pl);
    const long r = tt - k * pl;
    gy = surf_value(k, S.Gy);
    gz = 1 + (int)(r / S.Gx);
    gx = (int)(r - (long)(gz - 1) * S.Gx);
  } else {
    const long tt = t - S.cz - S.cy;
    const int niy = S.Gy - S.nsy, niz = S.Gz - S.nsz;
    const long pl = (long)niy * niz;
    const int k = (int)(tt / pl);
    const long r = tt - k * pl;
    gx = surf_value(k, S.Gx);
    gz = 1 + (int)(r / niy);
    gy = 1 + (int)(r - (long)(gz - 1) * niy);
  }
  dss_node_general<N>(G, B, recv, sl, gx, gy, gz);
}

template <int N>
__global__ __launch_bounds__(256) void k_surface(Geom G, Bound B, const double *__restrict__ recv,
                                                 Slots sl, SurfDesc S)
{
  const long total = S.cz + S.cy + S.cx;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x)
    surface_node<N>(G, B, recv, sl, S, t);
}

// The surface sums as extra blocks of another launch (mode 1 of the one-item
// kernel, whose item blocks leave every rank-surface location alone): blocks
// [sfirst, sfirst + nblk) walk the surface nodes, the item blocks start at
// block item0.
struct SurfWork {
  Bound B;
  const double *recv;
  Slots sl;
  SurfDesc S;
  unsigned nblk;    // surface blocks
  unsigned sfirst;  // first surface block (set at launch)
  unsigned item0;   // first item block (set at launch)
  int lead;         // surface blocks before the items (1) or after them (0)
};

// ---------------------------------------------------------------------------
// pack: the 26 messages (gpu/v-26-messages/Faces.cc:243-293)
// ---------------------------------------------------------------------------
// One launch, its blocks cut into ranges computed on the host: one range per
// face message and one for the 20 edge and corner messages.  A block of a
// face range copies kPackFaceEntries consecutive entries (y and z faces:
// runs of n or n^2 contiguous doubles in the field; x faces: the ix = 0 or
// n-1 column, doubles n apart, read straight from the field -- measured
// faster than staging whole elements in LDS to read them contiguously, 60^3
// n=7 8.0 us against 9.2, 40^3 n=12 11.4 against 14.8, equal for n <= 5);
// edge and corner entries find their slot among the 20 small ones.
// dst[k] is the arena offset entry 0 of slot k goes to: the send slot k, or
// -- fusing the self-halo exchange -- the receive slot opposite(k); -1 skips
// the slot.

constexpr int kPackFaceEntries = 1024;

struct PackDesc {
  long src_off[27];  // prefix sums of slot sizes (entries)
  long dst[26];
  int bstart[8];     // block prefix: ranges 0..5 the faces, 6 the edges and corners
};

// Field index of entry e of slot k (the reference's pack maps, Faces.cc:243-293).
template <int N>
__device__ __forceinline__ long pack_index(const Geom &G, int k, long e)
{
  const int n = nval<N>(G.n), nm1 = n - 1;
  int ix, iy, iz, jx, jy, jz;
  if (k < 6) {
    const int s = k & 1;
    const int a = (int)(e % n);
    const int b = (int)((e / n) % n);
    const long rest = e / ((long)n * n);
    if (k < 2) {  // x-face (iy,iz,jy,jz)
      iy = a; iz = b; jy = (int)(rest % G.my); jz = (int)(rest / G.my);
      ix = s ? nm1 : 0; jx = s ? G.mx - 1 : 0;
    } else if (k < 4) {  // y-face (ix,iz,jx,jz)
      ix = a; iz = b; jx = (int)(rest % G.mx); jz = (int)(rest / G.mx);
      iy = s ? nm1 : 0; jy = s ? G.my - 1 : 0;
    } else {  // z-face (ix,iy,jx,jy)
      ix = a; iy = b; jx = (int)(rest % G.mx); jy = (int)(rest / G.mx);
      iz = s ? nm1 : 0; jz = s ? G.mz - 1 : 0;
    }
  } else if (k < 18) {
    const int f = (k - 6) / 4, ee = (k - 6) % 4, a = ee & 1, b = ee >> 1;
    const int i = (int)(e % n);
    const int j = (int)(e / n);
    if (f == 0) {  // x-edge: (sy, sz) = (a, b)
      ix = i; jx = j;
      iy = a ? nm1 : 0; jy = a ? G.my - 1 : 0;
      iz = b ? nm1 : 0; jz = b ? G.mz - 1 : 0;
    } else if (f == 1) {  // y-edge: (sx, sz) = (a, b)
      iy = i; jy = j;
      ix = a ? nm1 : 0; jx = a ? G.mx - 1 : 0;
      iz = b ? nm1 : 0; jz = b ? G.mz - 1 : 0;
    } else {  // z-edge: (sx, sy) = (a, b)
      iz = i; jz = j;
      ix = a ? nm1 : 0; jx = a ? G.mx - 1 : 0;
      iy = b ? nm1 : 0; jy = b ? G.my - 1 : 0;
    }
  } else {
    const int c = k - 18, sx = c & 1, sy = (c >> 1) & 1, sz = c >> 2;
    ix = sx ? nm1 : 0; jx = sx ? G.mx - 1 : 0;
    iy = sy ? nm1 : 0; jy = sy ? G.my - 1 : 0;
    iz = sz ? nm1 : 0; jz = sz ? G.mz - 1 : 0;
  }
  const long n3 = (long)n * n * n;
  return ix + (long)n * (iy + (long)n * iz) + jx * n3 + jy * G.ys + jz * G.ps;
}

template <int N>
__global__ __launch_bounds__(256) void k_pack(Geom G, const PackDesc P, double *__restrict__ arena)
{
  const double *__restrict__ u = G.u;
  const int b = blockIdx.x, tid = threadIdx.x;
  int r = 0;
  while (r < 6 && b >= P.bstart[r + 1]) r++;  // block-uniform
  const int lb = b - P.bstart[r];
  if (r < 6) {
    const long dk = P.dst[r];
    if (dk < 0) return;
    const long cnt = P.src_off[r + 1] - P.src_off[r];
#pragma unroll
    for (int it = 0; it < kPackFaceEntries / 256; it++) {
      const long e = (long)lb * kPackFaceEntries + it * 256 + tid;
      if (e < cnt) arena[dk + e] = u[pack_index<N>(G, r, e)];
    }
    return;
  }
  // edges and corners
  const long t = P.src_off[6] + (long)lb * 256 + tid;
  if (t >= P.src_off[26]) return;
  int k = 6;
#pragma unroll 1
  while (k < 25 && t >= P.src_off[k + 1]) k++;
  const long dk = P.dst[k];
  if (dk < 0) return;
  const long e = t - P.src_off[k];
  arena[dk + e] = u[pack_index<N>(G, k, e)];
}

// Self-halo exchange: recv[k] <- send[opposite(k)] for every existing
// neighbour k (stand-in for the 26 MPI_Irecv/MPI_Isend pairs,
// gpu/v-26-messages/Faces.cc:144-171, 386-417).
struct XchgDesc {
  long idx_off[27];  // prefix sums over the copied slots (index space)
  long src[26];      // arena offset of send[opposite(k)]
  long dst[26];      // arena offset of recv[k]
  int nslots;
};

static __global__ __launch_bounds__(256) void k_exchange(const XchgDesc X, double *__restrict__ arena)
{
  const long total = X.idx_off[X.nslots];
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    int k = 0;
#pragma unroll 1
    while (k < X.nslots - 1 && t >= X.idx_off[k + 1]) k++;
    const long e = t - X.idx_off[k];
    arena[X.dst[k] + e] = arena[X.src[k] + e];
  }
}

// ---------------------------------------------------------------------------
// seeded fill (inputs of the benchmark configurations)
// ---------------------------------------------------------------------------

__host__ __device__ inline uint64_t hash64(uint64_t seed, uint64_t idx)
{
  uint64_t z = seed ^ (idx * 0x9E3779B97F4A7C15ull);
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// The 6-message schedule (gpu/v-6-messages-basic/Faces.cc:65-172): one phase
// per axis, z then y then x.  A lane owns one face pair across an element
// boundary along the axis -- the low face of element j and the high face of
// element j-1 (j = 0: the rank boundary pair, element m-1's high face, whose
// copies are the self-halo received faces) -- and writes both copies, so the
// phase runs in place without races.  Faces carry the sums of the earlier
// phases forward (edges and corners).  lo/hi: the low/high neighbour exists;
// compute = 0 sums internal pairs only on the rank surface.
// ---------------------------------------------------------------------------
static __global__ __launch_bounds__(256) void k_six(Geom G, int axis, int lo, int hi, int compute)
{
  const int n = G.n, nm1 = n - 1;
  const int m[3] = {G.mx, G.my, G.mz};
  const int b1 = (axis == 0) ? 1 : 0, b2 = (axis == 2) ? 1 : 2;
  const long total = (long)n * n * m[b1] * m[b2] * m[axis];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  long r = t;
  const int i1 = (int)(r % n); r /= n;
  const int i2 = (int)(r % n); r /= n;
  const int q1 = (int)(r % m[b1]); r /= m[b1];
  const int q2 = (int)(r % m[b2]); r /= m[b2];
  const int j = (int)r;
  const long es[3] = {(long)n * n * n, G.ys, G.ps};  // element strides
  const long ps_[3] = {1, n, (long)n * n};            // point strides
  const long base = q1 * es[b1] + q2 * es[b2] + i1 * ps_[b1] + i2 * ps_[b2];
  const long ia = base + (long)j * es[axis];
  const long ib = base + (long)(j == 0 ? m[axis] - 1 : j - 1) * es[axis] + (long)nm1 * ps_[axis];
  const double a = G.u[ia], b = G.u[ib];
  if (j > 0) {
    const bool surf = (q1 == 0 && i1 == 0) || (q1 == m[b1] - 1 && i1 == nm1) || (q2 == 0 && i2 == 0) ||
                      (q2 == m[b2] - 1 && i2 == nm1);
    if (compute || surf) {
      const double sum = a + b;
      G.u[ia] = sum;
      G.u[ib] = sum;
    }
  } else {
    if (lo) G.u[ia] = a + b;
    if (hi) G.u[ib] = b + a;
  }
}

// kind 0: integer-valued in [0,1023]; 1: uniform [-1,1); 2: the reference's
// positional init jz*1e4+jy*100+jx+iz*0.01+iy*1e-4+ix*1e-6 (gpu/main.cc:10-19),
// evaluated left to right without contraction.
static __global__ __launch_bounds__(256) void k_fill(Geom G, int kind, uint64_t seed)
{
  const int n = G.n;
  const long n3 = (long)n * n * n;
  const long total = n3 * G.mx * G.my * G.mz;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const long el = t / n3;
    const int r = (int)(t - el * n3);
    const int ix = r % n, iy = (r / n) % n, iz = r / (n * n);
    const int jx = (int)(el % G.mx);
    const int jy = (int)((el / G.mx) % G.my);
    const int jz = (int)(el / ((long)G.mx * G.my));
    double v;
    if (kind == 2) {
      v = __dadd_rn(__dadd_rn(__dadd_rn(__dadd_rn(__dadd_rn((double)(jz * 10000), (double)(jy * 100)),
                                                  (double)jx),
                                        __dmul_rn((double)iz, 0.01)),
                              __dmul_rn((double)iy, 0.0001)),
                    __dmul_rn((double)ix, 0.000001));
    } else {
      const uint64_t h = hash64(seed, (uint64_t)t);
      v = (kind == 0) ? (double)(h >> 54) : 2.0 * ((double)(h >> 11) * 0x1.0p-53) - 1.0;
    }
    G.u[r + jx * n3 + jy * G.ys + jz * G.ps] = v;
  }
}

}  // namespace faces
