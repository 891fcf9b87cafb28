// faces_stream.hpp -- the streaming in-place DSS kernel for periodic fields.
//
// Work item: S consecutive elements [jx0, jx0+S) of one element row (jy,jz),
// in memory order.  The item owns the lower-corner nodes (ix,iy,iz < n-1) of
// its elements; their copies form the item's footprint:
//   OWN  its elements, offsets [0, (n-1)*n^2) of each (planes iz < n-1; the
//        x-rows iy = n-1 in there are copies of the next row's nodes: loaded,
//        never stored)
//   Y    row (jy-1,jz):   x-rows iy = n-1 of planes iz < n-1
//   Z    row (jy,jz-1):   plane iz = n-1, x-rows iy < n-1
//   W    row (jy-1,jz-1): plane iz = n-1, x-row iy = n-1
//   SEAM the ix = n-1 column of element jx0-1 (m-1 when jx0 = 0) in the four
//        rows: the x partners of the first element's ix = 0 nodes
// and the ix = n-1 positions of the last element belong to the next item.
// Every field location is a copy of exactly one node, so each is loaded and
// stored by exactly one item: the kernel moves the field once in, once out.
//
// Cost model.  The footprint layout depends only on (n, S), both template
// parameters; an item changes nothing but four row bases.  Each thread owns
// fixed slots t = j*256 + tid of the footprint: OWN fills the first slots
// (a block-uniform base plus a per-thread offset computed once per launch),
// Y/Z/W/SEAM the rest (per-lane base picked once per item).  Loads and stores
// are therefore straight-line code with no per-access index arithmetic, and
// the compiler can wait for exactly the loads it needs (vmcnt on this
// architecture also counts the preceding item's stores).  The summation
// touches only the (n-1)^3 - (n-2)^3 shared nodes of each element, whose LDS
// positions are precomputed.  Blocks take work units (rows, column pieces,
// row groups) in XCD-contiguous order, several each (persistent blocks) or
// one each; a block issues an item's stores and then, at once, the next
// item's loads.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "faces_kernels.hpp"

namespace faces {

constexpr int kStreamThreads = 256;

// Buffer-addressed kernels use 32-bit byte offsets.  Offsets from kOobBase
// down to kOobBase - 8 * kOobSlots are sent out of range on purpose (loads and
// stores the hardware drops); a field, or in per-plane addressing one plane,
// is addressed this way only when its bytes stay below kMaxBufferBytes, so
// none of those offsets can reach it.
constexpr unsigned kOobBase = 0xFFFFFFF8u;
constexpr int kOobSlots = 64;
constexpr long kMaxBufferBytes = 0xFFFFF000L;
static_assert((long)kOobBase - 8L * kOobSlots >= kMaxBufferBytes, "dropped offsets must lie above every field");

template <int N, int S>
struct StreamShape {
  static constexpr int n = N, K = N - 1, n2 = N * N, n3 = N * N * N;
  static constexpr int cOWN = S * K * n2;
  static constexpr int JO = (cOWN + kStreamThreads - 1) / kStreamThreads;  // OWN slots
  static constexpr int cY = S * K * n, cZ = S * K * n, cW = S * n, cS = n2;
  static constexpr int cPART = cY + cZ + cW + cS;
  static constexpr int JP = (cPART + kStreamThreads - 1) / kStreamThreads;  // other slots
  static constexpr int J = JO + JP;
  static constexpr int P = J * kStreamThreads;  // LDS doubles
  static constexpr int rY = JO * kStreamThreads, rZ = rY + cY, rW = rZ + cZ, rS = rW + cW;
  static constexpr int nshared = K * K * K - (K - 1) * (K - 1) * (K - 1);
  static constexpr int NN = (S * nshared + kStreamThreads - 1) / kStreamThreads;
};

struct StreamItems {
  int my, mz;
  int nseg;             // mx / S
  int nitems;           // my*mz*nseg
  unsigned field_bytes; // below kMaxBufferBytes: offsets are 32-bit (BIG kernels: one plane's bytes)
  int xcd_remap;        // 1: contiguous ranges per XCD
  int walk;             // 0 items, 1 rows in x (seam carried), 2 columns in y (y rows carried),
                        // 3 groups of `rows` rows, segment by segment (seam and y rows carried)
  int rows;             // walk 3: rows per group; walk 2: rows per column piece (divides my)
  int skip_surface;     // 1: sum only nodes off the rank surface (open ranks:
                        // the surface is summed with the received copies);
                        // 2: open rank without neighbours: a rank-surface
                        // node sums only its copies on its own side of the
                        // rank surface (the wrap partners form nodes of their own)
  int nbr_mask;         // skip_surface = 2: face neighbours of the rank, bit 2a (2a+1) for
                        // the low (high) side of axis a: that side's surface nodes also
                        // sum the copies across the surface (stood in for by self halos)
};

// LDS doubles of the streaming kernel: the item footprint plus, for the walks
// that carry the seam, one saved seam column per row of the group.
template <int N, int S>
__host__ __device__ constexpr int stream_lds_doubles(int walk, int rows)
{
  return StreamShape<N, S>::P +
         ((walk == 1) ? 1 : (walk == 3) ? rows : 0) * StreamShape<N, S>::cS;
}

// BIG: the field spans 4 GB or more.  Each item then addresses its two
// planes (jz and jz-1, or the wrapped plane) through a descriptor of its own
// with 32-bit offsets inside the plane; slots of the partner region that may
// lie in either plane are loaded and stored once through each descriptor,
// the lanes of the other plane sent out of range.
// OPEN: an open rank (skip_surface = 2): a rank-surface node sums the copies
// on its own side of the surface, and those across it when the rank has a
// neighbour on that side, stood in for by a self halo (instantiated apart,
// so the periodic kernels keep their register allocation).
template <int N, int S, bool BIG = false, bool OPEN = false>
__global__ __launch_bounds__(kStreamThreads) void k_stream(Geom G, StreamItems I)
{
  using SH = StreamShape<N, S>;
  constexpr int n = N, K = N - 1, n2 = N * N;
  constexpr int J = SH::J, JO = SH::JO, NN = SH::NN;
  constexpr long n3 = (long)N * N * N;
  extern __shared__ double lds[];
  const int tid = threadIdx.x;

  // ---- per-thread footprint table (item-invariant)
  unsigned voff[J];
  unsigned keepm = 0;
  unsigned rowsel[SH::JP];  // PART slots: row 0 own, 1 (jy-1), 2 (jz-1), 3 both; bit 2 = seam
  // Walks carry part of the footprint instead of reloading it: the y-partner
  // rows come from the previous item's LDS (y walks), the seam column from a
  // copy saved in LDS by the same row's previous segment (x walks).
  const bool zig = I.walk == 1 || I.walk == 3;  // segment by segment along rows
  const int R = (I.walk == 3) ? I.rows : 1;     // rows per group
  unsigned carryY = 0, carryS = 0;  // PART slots carried: y-partner rows / seam
  unsigned carrysrc[SH::JP];        // LDS position, in the source item, of a carried value
  unsigned ringq[SH::JP];           // seam index (0..n^2) of a seam slot
#pragma unroll
  for (int j = 0; j < JO; j++) {
    const int q = j * kStreamThreads + tid;
    voff[j] = 0;
    if (q < SH::cOWN) {
      const int e = q / (K * n2), rr = q - e * (K * n2);
      const int ix = rr % n, iy = (rr / n) % n;
      voff[j] = e * n3 + rr;
      if (iy < K && !(e == S - 1 && ix == K)) keepm |= 1u << j;
    }
  }
#pragma unroll
  for (int jp = 0; jp < SH::JP; jp++) {
    const int j = JO + jp;
    int q = jp * kStreamThreads + tid;
    unsigned off = 0, rs = 0;
    bool keep = false;
    if (q < SH::cY) {
      const int e = q / (K * n), rr = q - e * (K * n);
      const int iz = rr / n, ix = rr - (rr / n) * n;
      off = e * n3 + iz * n2 + K * n + ix;
      rs = 1;
      keep = !(e == S - 1 && ix == K);
    } else if ((q -= SH::cY) < SH::cZ) {
      const int e = q / (K * n), rr = q - e * (K * n);
      const int ix = rr % n;
      off = e * n3 + K * n2 + rr;
      rs = 2;
      keep = !(e == S - 1 && ix == K);
    } else if ((q -= SH::cZ) < SH::cW) {
      const int e = q / n, ix = q - (q / n) * n;
      off = e * n3 + K * n2 + K * n + ix;
      rs = 3;
      keep = !(e == S - 1 && ix == K);
    } else if ((q -= SH::cW) < SH::cS) {
      keep = true;
      if (q < K * K) {
        off = (q / K) * n2 + (q % K) * n + K;
        rs = 4 | 0;
      } else if (q < K * K + K) {
        off = (q - K * K) * n2 + K * n + K;
        rs = 4 | 1;
      } else if (q < K * K + 2 * K) {
        off = K * n2 + (q - K * K - K) * n + K;
        rs = 4 | 2;
      } else {
        off = K * n2 + K * n + K;
        rs = 4 | 3;
      }
    }
    voff[j] = off;
    rowsel[jp] = rs;
    if (keep) keepm |= 1u << j;
    carrysrc[jp] = 0;
    ringq[jp] = (rs & 4) ? (unsigned)q : 0u;
    if (zig && (rs & 4)) {
      // x walk: the seam is the previous segment's last element's ix = n-1
      // column, which that item never changes.
      carryS |= 1u << jp;
      if (q < K * K) carrysrc[jp] = (S - 1) * (K * n2) + (q / K) * n2 + (q % K) * n + K;
      else if (q < K * K + K) carrysrc[jp] = SH::rY + (S - 1) * (K * n) + (q - K * K) * n + K;
      else if (q < K * K + 2 * K) carrysrc[jp] = SH::rZ + (S - 1) * (K * n) + (q - K * K - K) * n + K;
      else carrysrc[jp] = SH::rW + (S - 1) * n + K;
    } else if ((I.walk == 2 || I.walk == 3) && rs == 1) {
      // y walk: the y-partner rows are the previous row's x-rows iy = n-1,
      // loaded (never changed) by that item as part of its OWN region.
      const int qy = jp * kStreamThreads + tid;
      const int e = qy / (K * n), rr = qy - e * (K * n);
      const int iz = rr / n, ix = rr - (rr / n) * n;
      carryY |= 1u << jp;
      carrysrc[jp] = e * (K * n2) + iz * n2 + K * n + ix;
    }
  }

  // ---- per-thread shared-node table: 8 LDS positions (16 bits each)
  unsigned np[NN][4];
  unsigned nflag[NN];  // bit0 cx, bit1 cy, bit2 cz, bit3 active
#pragma unroll
  for (int i = 0; i < NN; i++) {
    const int k = i * kStreamThreads + tid;
    nflag[i] = 0;
    np[i][0] = np[i][1] = np[i][2] = np[i][3] = 0;
    if (k >= S * SH::nshared) continue;
    const int e = k / SH::nshared;
    int r = k - e * SH::nshared;
    int ix, iy, iz;
    if (r < K * K) {
      iz = 0; iy = r / K; ix = r - (r / K) * K;
    } else if (r < K * K + K * (K - 1)) {
      r -= K * K;
      iz = 1 + r / K; iy = 0; ix = r - (r / K) * K;
    } else {
      r -= K * K + K * (K - 1);
      constexpr int K1 = (K > 1) ? K - 1 : 1;  // this branch needs K > 1
      iz = 1 + r / K1; iy = 1 + r - (r / K1) * K1; ix = 0;
    }
    unsigned p[8];
    p[0] = e * (K * n2) + iz * n2 + iy * n + ix;  // own
    p[2] = SH::rY + e * (K * n) + iz * n + ix;    // y partner
    p[4] = SH::rZ + e * (K * n) + iy * n + ix;    // z partner
    p[6] = SH::rW + e * n + ix;                   // yz partner
    if (e > 0) {
      p[1] = (e - 1) * (K * n2) + iz * n2 + iy * n + K;
      p[3] = SH::rY + (e - 1) * (K * n) + iz * n + K;
      p[5] = SH::rZ + (e - 1) * (K * n) + iy * n + K;
      p[7] = SH::rW + (e - 1) * n + K;
    } else {
      p[1] = SH::rS + iz * K + iy;
      p[3] = SH::rS + K * K + iz;
      p[5] = SH::rS + K * K + K + iy;
      p[7] = SH::rS + K * K + 2 * K;
    }
#pragma unroll
    for (int w = 0; w < 4; w++) np[i][w] = p[2 * w] | (p[2 * w + 1] << 16);
    nflag[i] = (ix == 0 ? 1u : 0u) | (iy == 0 ? 2u : 0u) | (iz == 0 ? 4u : 0u) | 8u | (e == 0 ? 16u : 0u);
  }

  // ---- traversal.  Work units are dealt to blocks in XCD-contiguous ranges
  // (block b runs on XCD b % 8) or in one global order.  A unit is one item
  // (walk 0), an element row walked in x (walk 1), L consecutive rows of
  // the column at fixed (jz, segment) walked in y (walk 2), or a group of R rows walked
  // segment by segment, the R rows of a segment in turn (walk 3).  Items that
  // share memory sectors then run back to back on one CU, whose L2 merges
  // their writes, and carried values are not fetched twice.
  const int L = (I.walk == 2) ? I.rows : 1;  // walk 2: rows per column piece
  const int units = (I.walk == 1) ? I.my * I.mz
                    : (I.walk == 2) ? I.mz * I.nseg * (I.my / L)
                    : (I.walk == 3) ? I.mz * (I.my / R)
                                    : I.nitems;
  int start, cnt, k, nbx;
  if (I.xcd_remap) {
    const int xcd = blockIdx.x & 7;
    nbx = (gridDim.x - xcd + 7) >> 3;  // blocks on this XCD
    const int q8 = units >> 3, rem8 = units & 7;
    start = xcd * q8 + (xcd < rem8 ? xcd : rem8);
    cnt = q8 + (xcd < rem8 ? 1 : 0);
    k = blockIdx.x >> 3;
  } else {
    start = 0;
    cnt = units;
    k = blockIdx.x;
    nbx = gridDim.x;
  }
  if (k >= cnt) return;
  const int nsegu = zig ? R * I.nseg : (I.walk == 2) ? L : 1;  // items per unit
  int sg = 0;
  auto item_of = [&](int kk, int ss) {
    const int c = start + kk;
    if (zig) return (c * R + ss % R) * I.nseg + ss / R;  // row c*R + r, segment ss/R
    // walk 2: c = (jz * (my / L) + piece) * nseg + segment; its rows start at
    // jz * my + piece * L = (c / nseg) * L
    if (I.walk == 2) return ((c / I.nseg) * L + ss) * I.nseg + (c % I.nseg);
    return c;
  };
  double *const ring = lds + SH::P;  // saved seam columns, one per row of the group

  // Byte offsets into the field through one buffer resource: a store whose
  // offset is out of range is dropped by the hardware, so the stores of
  // positions this item does not own need no branch, every memory
  // instruction is unconditional, and the compiler waits on exact counts.
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc(G.u, 0, (int)I.field_bytes, 0x00020000);
  constexpr unsigned kOOB = kOobBase;  // any offset at or above field_bytes is dropped
#pragma unroll
  for (int j = 0; j < J; j++) voff[j] *= 8;  // bytes from here on
  unsigned dropm[J];
#pragma unroll
  for (int j = 0; j < J; j++) dropm[j] = (keepm & (1u << j)) ? 0u : kOOB;

  unsigned obase = 0;          // item's OWN base (bytes, block-uniform)
  unsigned pbase[SH::JP];      // item's per-lane PART bases (bytes)
  // BIG: descriptors of the loaded item's planes jz (A) and jz-1 (B), and
  // the PART slots whose region lies in plane B
  __amdgpu_buffer_rsrc_t rA = rsrc, rB = rsrc;
  unsigned inB = 0;
#pragma unroll
  for (int jp = 0; jp < SH::JP; jp++)
    if ((rowsel[jp] & 3) >= 2) inB |= 1u << jp;
  auto bases = [&](int it) {
    const int sg = it % I.nseg;
    const int row = it / I.nseg;
    const int jy = row % G.my, jz = row / G.my;
    const int jym = (jy == 0) ? G.my - 1 : jy - 1;
    const int jzm = (jz == 0) ? G.mz - 1 : jz - 1;
    const int jx0 = sg * S;
    const int jxs = (jx0 == 0) ? G.mx - 1 : jx0 - 1;
    const long pa = BIG ? 0 : jz * G.ps, pb = BIG ? 0 : jzm * G.ps;
    const unsigned r0 = (unsigned)((pa + jy * G.ys) * 8);
    const unsigned r1 = (unsigned)((pa + jym * G.ys) * 8);
    const unsigned r2 = (unsigned)((pb + jy * G.ys) * 8);
    const unsigned r3 = (unsigned)((pb + jym * G.ys) * 8);
    if constexpr (BIG) {
      rA = __builtin_amdgcn_make_buffer_rsrc(G.u + jz * G.ps, 0, (int)I.field_bytes, 0x00020000);
      rB = __builtin_amdgcn_make_buffer_rsrc(G.u + jzm * G.ps, 0, (int)I.field_bytes, 0x00020000);
    }
    obase = r0 + (unsigned)(jx0 * n3 * 8);
#pragma unroll
    for (int jp = 0; jp < SH::JP; jp++) {
      const unsigned rs = rowsel[jp];
      const unsigned rb = ((rs & 3) == 0) ? r0 : ((rs & 3) == 1) ? r1 : ((rs & 3) == 2) ? r2 : r3;
      pbase[jp] = rb + (unsigned)(((rs & 4) ? jxs : jx0) * n3 * 8);
    }
  };
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  double v[J];
  // A carried slot is not loaded: its load is sent out of range (no memory
  // access, the instruction count is unchanged).
  auto load_item = [&](bool cy, bool cs) {
    const unsigned cm = (cy ? carryY : 0u) | (cs ? carryS : 0u);
#pragma unroll
    for (int j = 0; j < J; j++) {
      unsigned off = ((j < JO) ? obase : pbase[j - JO < 0 ? 0 : j - JO]) + voff[j];
      if (j >= JO && (cm & (1u << (j - JO)))) off = kOOB;
      if constexpr (!BIG) {
        v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0));
      } else if (j < JO) {
        v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rA, off, 0, 0));
      } else {
        const bool b = inB & (1u << (j - JO < 0 ? 0 : j - JO));
        const double xa =
            __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rA, b ? kOOB : off, 0, 0));
        const double xb =
            __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rB, b ? off : kOOB, 0, 0));
        v[j] = b ? xb : xa;
      }
    }
  };

  int it = item_of(k, 0);
  bases(it);
  load_item(false, false);
  bool cy = false, cs = false;  // the current item's y rows / seam are carried
  double vcar[SH::JP];
#pragma unroll
  for (int jp = 0; jp < SH::JP; jp++) vcar[jp] = 0.0;
  // Dropped stores (offset out of range), as many as an item issues: the loop
  // is entered with the same pattern of outstanding operations as every later
  // iteration -- an item's loads, then its stores -- so the waits at the loop
  // head are exact.
  constexpr int NST = J + (BIG ? SH::JP : 0);
  static_assert(NST <= kOobSlots, "dropped stores below the reserved offsets");
#pragma unroll
  for (int j = 0; j < NST; j++)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, 0.0), rsrc, kOOB - 8 * j, 0, 0);
  __amdgpu_buffer_rsrc_t cA = rA, cB = rB;  // the current item's descriptors (BIG)
  while (true) {
    __syncthreads();  // the previous item's stores have read LDS
    const int rr = sg % R;  // the current item's row within the group
#pragma unroll
    for (int j = 0; j < J; j++) {
      double x = v[j];
      if (j >= JO) {
        const int jp = j - JO < 0 ? 0 : j - JO;
        if (cy && (carryY & (1u << jp))) x = vcar[jp];
        if (cs && (carryS & (1u << jp))) x = ring[rr * SH::cS + ringq[jp]];
      }
      lds[j * kStreamThreads + tid] = x;
    }
    // this item's store offsets, then the next item's loads (in flight
    // during this item's sums and stores)
    unsigned soff[J];
#pragma unroll
    for (int j = 0; j < J; j++)
      soff[j] = (((j < JO) ? obase : pbase[j - JO < 0 ? 0 : j - JO]) + voff[j]) | dropm[j];
    int nk = k, ns = sg + 1;
    if (ns == nsegu) {
      ns = 0;
      nk = k + nbx;
    }
    const bool more = nk < cnt;
    // the next item's carried parts (only within one unit)
    const bool ncy = more && ns > 0 && (I.walk == 2 || (I.walk == 3 && ns % R != 0));
    const bool ncs = more && zig && ns >= R;
    // Unconditional loads keep the wait accounting exact; after the last item
    // they re-read its own locations (in bounds, never used).
    const int nit = more ? item_of(nk, ns) : it;
    cA = rA;
    cB = rB;
    bases(nit);
    load_item(ncy, ncs);
    __syncthreads();
    if (ncy) {
#pragma unroll
      for (int jp = 0; jp < SH::JP; jp++)
        if (carryY & (1u << jp)) vcar[jp] = lds[carrysrc[jp]];
    }
    if (zig && sg + R < nsegu) {  // the row's next segment takes this seam
#pragma unroll
      for (int jp = 0; jp < SH::JP; jp++)
        if (carryS & (1u << jp)) ring[rr * SH::cS + ringq[jp]] = lds[carrysrc[jp]];
    }
    // shared nodes: x pairs, then y, then z
    unsigned skip_seam = 0;  // axes on which this item holds the rank surface
    if (I.skip_surface) {
      const int row = it / I.nseg;
      skip_seam = ((it % I.nseg) == 0 ? 1u : 0u) | ((row % G.my) == 0 ? 2u : 0u) |
                  ((row / G.my) == 0 ? 4u : 0u);
    }
#pragma unroll
    for (int i = 0; i < NN; i++) {
      const unsigned f = nflag[i];
      if (!(f & 8u)) continue;
      // Axes on which this node's pair is a periodic wrap pair (global
      // coordinate 0: the rank surface).  Skipping, such nodes are left to the
      // surface kernel; open without neighbours, the copies on each side of
      // those axes are summed apart (x pairs, then y, then z, over the other
      // shared axes), each side being a node of its own.
      const unsigned wrap = skip_seam & ((f & 6u) | ((f & 17u) == 17u ? 1u : 0u));
      if (wrap) continue;  // rank-surface node: the surface kernel's, or the pass below
      const unsigned p0 = np[i][0] & 0xffff, p1 = np[i][0] >> 16;
      const unsigned p2 = np[i][1] & 0xffff, p3 = np[i][1] >> 16;
      const unsigned p4 = np[i][2] & 0xffff, p5 = np[i][2] >> 16;
      const unsigned p6 = np[i][3] & 0xffff, p7 = np[i][3] >> 16;
      const bool cx = f & 1u, cy = f & 2u, cz = f & 4u;
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
    // Open rank: the rank-surface nodes of this item (block-uniform: only
    // items on the surface have any).  Each side of the surface is a node of
    // its own: it sums its copies, and those across the surface when the rank
    // has a neighbour on that side, x pairs, then y, then z.
    if constexpr (OPEN) {
      if (skip_seam) {
#pragma unroll
        for (int i = 0; i < NN; i++) {
          const unsigned f = nflag[i];
          if (!(f & 8u)) continue;
          const unsigned wrap = skip_seam & ((f & 6u) | ((f & 17u) == 17u ? 1u : 0u));
          if (!wrap) continue;
          // copy (sx | sy << 1 | sz << 2), side 1 = the partner across that
          // axis; all eight are read at once (positions of copies the node
          // does not have lie in the footprint too, and go unused)
          const unsigned pp[8] = {np[i][0] & 0xffff, np[i][0] >> 16, np[i][1] & 0xffff, np[i][1] >> 16,
                                  np[i][2] & 0xffff, np[i][2] >> 16, np[i][3] & 0xffff, np[i][3] >> 16};
          double v[8];
#pragma unroll
          for (int m = 0; m < 8; m++) v[m] = lds[pp[m]];
          // pair[a][side]: that side of axis a sums both copies -- a pair
          // inside the rank, or across the surface with a neighbour there
          const unsigned nb = (unsigned)I.nbr_mask;
          bool pair[3][2];
#pragma unroll
          for (int a = 0; a < 3; a++)
#pragma unroll
            for (int sd = 0; sd < 2; sd++)
              pair[a][sd] = (f & (1u << a)) && (!(wrap & (1u << a)) || (nb & (1u << (2 * a + sd))));
          // the tree level by level, for both sides of every axis: x pairs
          // (or the side's copy), then y, then z
          double A[2][2][2], Q[2][2][2], R[2][2][2];  // [gx][..][..]
#pragma unroll
          for (int gx = 0; gx < 2; gx++)
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const double x0 = v[2 * k], x1 = v[2 * k + 1];
              A[gx][k & 1][k >> 1] = pair[0][gx] ? x0 + x1 : (gx ? x1 : x0);
            }
#pragma unroll
          for (int gx = 0; gx < 2; gx++)
#pragma unroll
            for (int gy = 0; gy < 2; gy++)
#pragma unroll
              for (int az = 0; az < 2; az++)
                Q[gx][gy][az] = pair[1][gy] ? A[gx][0][az] + A[gx][1][az] : A[gx][gy][az];
#pragma unroll
          for (int gx = 0; gx < 2; gx++)
#pragma unroll
            for (int gy = 0; gy < 2; gy++)
#pragma unroll
              for (int gz = 0; gz < 2; gz++)
                R[gx][gy][gz] = pair[2][gz] ? Q[gx][gy][0] + Q[gx][gy][1] : Q[gx][gy][gz];
#pragma unroll
          for (int m = 0; m < 8; m++) {
            const bool ex = (!(m & 1) || (f & 1u)) && (!(m & 2) || (f & 2u)) && (!(m & 4) || (f & 4u));
            if (!ex) continue;
            // copy m's side on each split axis, picked with selects
            const bool gx = (wrap & 1u) && (m & 1), gy = (wrap & 2u) && (m & 2), gz = (wrap & 4u) && (m & 4);
            double rz[2][2];
#pragma unroll
            for (int x = 0; x < 2; x++)
#pragma unroll
              for (int y = 0; y < 2; y++) rz[x][y] = gz ? R[x][y][1] : R[x][y][0];
            const double ry0 = gy ? rz[0][1] : rz[0][0], ry1 = gy ? rz[1][1] : rz[1][0];
            lds[pp[m]] = gx ? ry1 : ry0;
          }
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < J; j++) {
      const u32x2 x = __builtin_bit_cast(u32x2, lds[j * kStreamThreads + tid]);
      if constexpr (!BIG) {
        __builtin_amdgcn_raw_buffer_store_b64(x, rsrc, soff[j], 0, 0);
      } else if (j < JO) {
        __builtin_amdgcn_raw_buffer_store_b64(x, cA, soff[j], 0, 0);
      } else {
        const bool b = inB & (1u << (j - JO < 0 ? 0 : j - JO));
        __builtin_amdgcn_raw_buffer_store_b64(x, cA, b ? kOOB : soff[j], 0, 0);
        __builtin_amdgcn_raw_buffer_store_b64(x, cB, b ? soff[j] : kOOB, 0, 0);
      }
    }
    if (!more) break;
    k = nk;
    sg = ns;
    it = nit;
    cy = ncy;
    cs = ncs;
  }
}

}  // namespace faces
