// faces_flow.hpp -- the one-item-per-block DSS kernel for periodic fields.
//
// The work item and its footprint are the streaming kernel's (faces_stream.hpp:
// S consecutive elements of one element row; OWN planes iz < n-1 of its
// elements, the Y, Z and W partner rows, the x SEAM column), laid out in LDS
// as StreamShape<N,S> describes, and every field location is loaded and
// stored by exactly one item.  What differs is the traversal: every block
// takes one item, and the items run in memory order -- XCD x (the hardware
// dispatches block b to XCD b % 8) takes the x-th contiguous eighth of the
// item list, and its blocks are dispatched in order along it.  At any time
// the blocks of an XCD work on a compact stretch of the field: an item's
// x-seam partner (the previous item) and y-partner row (nseg items back) were
// read and written moments before by the same XCD, so their lines meet again
// in that XCD's L2, where the partial-line writes of the two items merge
// before memory sees them.  A block holds no state beyond its item: no
// carried values, no dummy operations, no loop; short-lived blocks keep many
// items' loads in flight per CU (up to 8 blocks, set by the registers and
// LDS a block takes), as a plain streaming copy does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "faces_stream.hpp"

namespace faces {

struct FlowItems {
  int my, mz;
  int nseg;              // items per element row: the launch covers columns [jxb, jxb + nseg*S)
  int jxb;               // first element column of the launch's items
  int nitems;            // my*mz*nseg
  unsigned field_bytes;  // below kMaxBufferBytes: offsets are 32-bit
  int nbr_mask;          // MODE 2: face neighbours of the rank, bit 2a (2a+1) for the low (high)
                         // side of axis a (stood in for by self halos)
  int xcd_remap;         // 2: XCD x takes the x-th eighth of every plane's items, planes in
                         // order; 1: XCD x takes the x-th contiguous eighth of the list;
                         // 0: block b takes item b
};

// MODE 0: periodic (every node with its wrap partners).  MODE 1: the staged
// step's interior sums (faces_inner): nodes on the rank surface -- those
// whose partner across a periodic wrap the item holds -- are left as they
// are, for the surface sums that add the received copies.  MODE 2: an open
// rank: each side of the rank surface is a node of its own, which sums its
// copies on that side, and those across the surface where the rank has a
// neighbour there (stood in for by the self halo), x pairs, then y, then z.
// MODE 3: the staged step's interior and surface sums in one pass: as MODE
// 2, except that the copies across the rank surface are the received ones
// (the message arena W.recv, its receive slots W.sl), not the field's own
// copies on the other side: each side of the surface sums its local copies
// and, where the rank has a neighbour on that side, the copies that
// neighbour sent, in the tree order of the surface sums (faces_kernels.hpp,
// dss_node_general).  Compiled with LW 2.
// LW (load and store widths of the OWN region; the partner regions are moved
// 8 bytes per lane):
//   0  8 bytes per lane, loads and stores.
//   1  16-byte loads on 16-byte boundaries (the field must be 16-byte
//      aligned): an element's run [b, b + (n-1)n^2) has even length; it takes
//      (n-1)n^2/2 + 1 pair slots, of which an even b uses the first
//      (n-1)n^2/2 and an odd b all, starting one double early and ending one
//      late (the halves outside the run are dropped on the way to LDS);
//      stores 8 bytes per lane, since a pair may hold one location the item
//      owns and one it must not write.
//   2  16-byte loads and stores of exactly the locations the item owns,
//      which form runs -- per element but the last, one run of (n-1)n
//      doubles per plane iz < n-1 (its x-rows iy < n-1); per last element,
//      one run of n-1 per x-row (ix < n-1) -- split into pairs from their
//      starts, so a pair may start on any double (8-byte aligned); a run of
//      odd length ends in a single moved 8 bytes per lane.  The rows iy = n-1
//      and the last element's column ix = n-1, which the item neither sums
//      nor stores, are not loaded.
// TREE (the order of a node's sum tree): 0 x pairs, then y, then z (the
// 26-message step); 1 z pairs, then y, then x (the 6-message schedule's three
// phases, gpu/v-6-messages-basic/Faces.cc:65-172, whose z phase sums the z
// pairs, its y phase the pairs of those sums, its x phase the rest).  Tree 1
// is tree 0 with the x and z axes exchanged: the copies and flags of a node
// are relabelled (copy bits sx <-> sz) before the same code sums them.
// LDS words of the MODE 1 marks (one bit per footprint position), placed
// after the footprint's StreamShape<N, S>::P doubles.
template <int N, int S>
constexpr int kFlowMarkWords = (StreamShape<N, S>::P + 31) / 32;

template <int N, int S, int LW, int MODE = 0, int TREE = 0>
__global__ __launch_bounds__(kStreamThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_flow(Geom G, FlowItems I, const SurfWork W)
{
  using SH = StreamShape<N, S>;
  constexpr int n = N, K = N - 1, n2 = N * N;
  constexpr unsigned n3 = (unsigned)(N * N * N);
  constexpr int J = SH::J, JO = SH::JO, JP = SH::JP, NN = SH::NN;
  constexpr unsigned kOOB = kOobBase;
  extern __shared__ double lds[];
  const int tid = threadIdx.x;

  // MODE 1: the surface blocks of the launch sum the rank surface meanwhile
  unsigned bid = blockIdx.x;
  if constexpr (MODE == 1) {
    if (bid - W.sfirst < W.nblk) {
      const long total = W.S.cz + W.S.cy + W.S.cx;
      const long stride = (long)W.nblk * kStreamThreads;
      for (long t = (long)(bid - W.sfirst) * kStreamThreads + tid; t < total; t += stride)
        surface_node<N>(G, W.B, W.recv, W.sl, W.S, t);
      return;
    }
    bid -= W.item0;
  }

  // ---- the item
  int it = bid;
  if (I.xcd_remap == 2) {
    // The x-th eighth of the items of every plane, planes in order: an
    // item's z partner is one plane's share (an eighth of a plane) back in
    // this XCD's walk, its y partner nseg items back.
    const int xcd = bid & 7, k = bid >> 3;
    const int per = G.my * I.nseg;  // items per plane
    const int i0 = xcd * per / 8, cnt = (xcd + 1) * per / 8 - i0;
    if (cnt == 0 || k >= cnt * G.mz) return;
    const int jzk = k / cnt;
    it = jzk * per + i0 + (k - jzk * cnt);
  } else if (I.xcd_remap == 1) {
    const int xcd = bid & 7, k = bid >> 3;
    const int q8 = I.nitems >> 3, rem8 = I.nitems & 7;
    if (k >= q8 + (xcd < rem8 ? 1 : 0)) return;
    it = xcd * q8 + (xcd < rem8 ? xcd : rem8) + k;
  }
  const int sg = it % I.nseg, row = it / I.nseg;
  const int jy = row % G.my, jz = row / G.my;
  const int jym = (jy == 0) ? G.my - 1 : jy - 1;
  const int jzm = (jz == 0) ? G.mz - 1 : jz - 1;
  const int jx0 = I.jxb + sg * S;
  const int jxs = (jx0 == 0) ? G.mx - 1 : jx0 - 1;
  // byte offsets of the four rows: own, y partner, z partner, yz partner
  const unsigned rb[4] = {(unsigned)((jz * G.ps + jy * G.ys) * 8), (unsigned)((jz * G.ps + jym * G.ys) * 8),
                          (unsigned)((jzm * G.ps + jy * G.ys) * 8), (unsigned)((jzm * G.ps + jym * G.ys) * 8)};
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(G.u, 0, (int)I.field_bytes, 0x00020000);

  // ---- footprint slots of this thread: load offset, and whether it stores
  unsigned off[J];
  unsigned keep = 0;
#pragma unroll
  for (int j = 0; j < JO; j++) {
    const int q = j * kStreamThreads + tid;
    off[j] = kOOB;
    if (q < SH::cOWN) {
      const int e = q / (K * n2), rr = q - e * (K * n2);
      const int ix = rr % n, iy = (rr / n) % n;
      off[j] = rb[0] + ((unsigned)(jx0 + e) * n3 + (unsigned)rr) * 8;
      // x-rows iy = n-1 are the next row's nodes; the last element's ix = n-1
      // column is the next item's
      if (iy < K && !(e == S - 1 && ix == K)) keep |= 1u << j;
    }
  }
#pragma unroll
  for (int jp = 0; jp < JP; jp++) {
    const int j = JO + jp;
    int q = jp * kStreamThreads + tid;
    unsigned o = kOOB;
    bool kp = false;
    if (q < SH::cY) {  // row jy-1: x-rows iy = n-1 of planes iz < n-1
      const int e = q / (K * n), rr = q - e * (K * n);
      const int iz = rr / n, ix = rr - (rr / n) * n;
      o = rb[1] + ((unsigned)(jx0 + e) * n3 + (unsigned)(iz * n2 + K * n + ix)) * 8;
      kp = !(e == S - 1 && ix == K);
    } else if ((q -= SH::cY) < SH::cZ) {  // row (jy, jz-1): plane iz = n-1, x-rows iy < n-1
      const int e = q / (K * n), rr = q - e * (K * n);
      const int ix = rr % n;
      o = rb[2] + ((unsigned)(jx0 + e) * n3 + (unsigned)(K * n2 + rr)) * 8;
      kp = !(e == S - 1 && ix == K);
    } else if ((q -= SH::cZ) < SH::cW) {  // row (jy-1, jz-1): plane iz = n-1, x-row iy = n-1
      const int e = q / n, ix = q - (q / n) * n;
      o = rb[3] + ((unsigned)(jx0 + e) * n3 + (unsigned)(K * n2 + K * n + ix)) * 8;
      kp = !(e == S - 1 && ix == K);
    } else if ((q -= SH::cW) < SH::cS) {  // the seam: ix = n-1 of element jx0-1 in the four rows
      kp = true;
      const unsigned sb = (unsigned)jxs * n3 * 8;
      if (q < K * K) o = rb[0] + sb + (unsigned)((q / K) * n2 + (q % K) * n + K) * 8;
      else if (q < K * K + K) o = rb[1] + sb + (unsigned)((q - K * K) * n2 + K * n + K) * 8;
      else if (q < K * K + 2 * K) o = rb[2] + sb + (unsigned)(K * n2 + (q - K * K - K) * n + K) * 8;
      else o = rb[3] + sb + (unsigned)(K * n2 + K * n + K) * 8;
    }
    off[j] = o;
    if (kp) keep |= 1u << j;
  }

  // ---- loads of the whole footprint, all in flight at once
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  constexpr int L = K * n2;               // OWN run of one element (even)
  constexpr int PE = L / 2 + 1;           // pair slots per element (W16)
  constexpr int JW = LW == 1 ? (S * PE + kStreamThreads - 1) / kStreamThreads : 1;
  double v[J];
  u32x4 w[JW];
  unsigned wq[JW];  // LW 1: LDS position of the pair's high half | bit 16: low valid | bit 17: high valid
  // LW 2: pair slots over the owned runs, and singles ending odd runs
  constexpr int RP = (K * n) / 2;                          // pairs per plane run
  constexpr int PNL = K * RP;                              // pairs of an element but the last
  constexpr int PRW = K / 2, PRW1 = PRW > 0 ? PRW : 1;    // pairs per row run of the last element
  constexpr int NPR = (S - 1) * PNL + K * K * PRW;
  constexpr int NSG = (K & 1) ? K * K : 0;                 // singles of the last element (n even)
  // LW 3: the partner regions' owned runs in pairs too (z: an element's x-rows
  // iy < n-1 of plane n-1; y: the x-row iy = n-1 of each plane iz < n-1; yz:
  // the x-row iy = n-1 of plane n-1; the last element's runs end at ix = n-2),
  // odd runs ending in a single, and the seam as singles
  constexpr int ZPF = (K * n) / 2, ZPL = K / 2, ZPL1 = ZPL > 0 ? ZPL : 1;
  constexpr int XPF = n / 2, XPL = K / 2, XPL1 = XPL > 0 ? XPL : 1;  // a run of one x-row
  constexpr int NZP = (S - 1) * ZPF + K * ZPL;
  constexpr int NYP = (S - 1) * K * XPF + K * XPL;
  constexpr int NWP = (S - 1) * XPF + XPL;
  constexpr int NZS = (K & 1) ? K : 0;                              // last element's z rows
  constexpr int NYS = (S - 1) * K * (n & 1) + ((K & 1) ? K : 0);
  constexpr int NWS = (S - 1) * (n & 1) + (K & 1);
  constexpr int NPP = LW == 3 ? NZP + NYP + NWP : 0;
  constexpr int NPS = LW == 3 ? NZS + NYS + NWS + SH::cS : 0;
  constexpr int JR = (LW >= 2 && NPR + NPP > 0) ? (NPR + NPP + kStreamThreads - 1) / kStreamThreads : 1;
  constexpr int JG = (LW >= 2 && NSG + NPS > 0) ? (NSG + NPS + kStreamThreads - 1) / kStreamThreads : 0;
  unsigned roff[JR], rq[JR];             // byte offset (kOOB: none), LDS position of the low half
  unsigned goff[JG > 0 ? JG : 1], gq[JG > 0 ? JG : 1];
  u32x4 wr[JR];
  double wg[JG > 0 ? JG : 1];
  if constexpr (LW >= 2) {
#pragma unroll
    for (int jr = 0; jr < JR; jr++) {
      const int t = jr * kStreamThreads + tid;
      roff[jr] = kOOB;
      rq[jr] = 0;
      if (LW == 3 && t >= NPR && t < NPR + NPP) {  // a pair of a partner region
        int tp = t - NPR;
        int e, ro, lq, rbi;  // element, offset in the element, LDS position, row
        if (tp < NZP) {
          rbi = 2;
          if (tp < (S - 1) * ZPF) {
            e = tp / ZPF;
            const int r = 2 * (tp - e * ZPF);
            ro = K * n2 + r;
            lq = SH::rZ + e * (K * n) + r;
          } else {
            const int t2 = tp - (S - 1) * ZPF, iy = t2 / ZPL1, ix = 2 * (t2 - iy * ZPL1);
            e = S - 1;
            ro = K * n2 + iy * n + ix;
            lq = SH::rZ + e * (K * n) + iy * n + ix;
          }
        } else if ((tp -= NZP) < NYP) {
          rbi = 1;
          int iz, ix;
          if (tp < (S - 1) * K * XPF) {
            const int run = tp / XPF;
            e = run / K;
            iz = run - e * K;
            ix = 2 * (tp - run * XPF);
          } else {
            const int t2 = tp - (S - 1) * K * XPF;
            e = S - 1;
            iz = t2 / XPL1;
            ix = 2 * (t2 - iz * XPL1);
          }
          ro = iz * n2 + K * n + ix;
          lq = SH::rY + e * (K * n) + iz * n + ix;
        } else {
          tp -= NYP;
          rbi = 3;
          int ix;
          if (tp < (S - 1) * XPF) {
            e = tp / XPF;
            ix = 2 * (tp - e * XPF);
          } else {
            e = S - 1;
            ix = 2 * (tp - (S - 1) * XPF);
          }
          ro = K * n2 + K * n + ix;
          lq = SH::rW + e * n + ix;
        }
        roff[jr] = rb[rbi] + ((unsigned)(jx0 + e) * n3 + (unsigned)ro) * 8;
        rq[jr] = (unsigned)lq;
      }
      if (t < NPR) {
        int e, r0;
        if (t < (S - 1) * PNL) {
          e = t / PNL;
          const int rem = t - e * PNL, iz = rem / RP;
          r0 = iz * n2 + 2 * (rem - iz * RP);
        } else {
          const int t2 = t - (S - 1) * PNL, row = t2 / PRW1, iz = row / K, iy = row - iz * K;
          e = S - 1;
          r0 = iz * n2 + iy * n + 2 * (t2 - row * PRW1);
        }
        roff[jr] = rb[0] + ((unsigned)(jx0 + e) * n3 + (unsigned)r0) * 8;
        rq[jr] = (unsigned)(e * (K * n2) + r0);
      }
      wr[jr] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, roff[jr], 0, 0);
    }
#pragma unroll
    for (int jg = 0; jg < JG; jg++) {
      const int t = jg * kStreamThreads + tid;
      goff[jg] = kOOB;
      gq[jg] = 0;
      if (t < NSG) {
        const int iz = t / K, iy = t - (t / K) * K, r = iz * n2 + iy * n + (K - 1);
        goff[jg] = rb[0] + ((unsigned)(jx0 + S - 1) * n3 + (unsigned)r) * 8;
        gq[jg] = (unsigned)((S - 1) * (K * n2) + r);
      } else if (LW == 3 && t < NSG + NPS) {  // a single of a partner region, or the seam
        int u = t - NSG;
        if (u < NZS) {  // the last element's z rows (odd length)
          const int ro = K * n2 + u * n + (K - 1);
          goff[jg] = rb[2] + ((unsigned)(jx0 + S - 1) * n3 + (unsigned)ro) * 8;
          gq[jg] = (unsigned)(SH::rZ + (S - 1) * (K * n) + u * n + (K - 1));
        } else if ((u -= NZS) < NYS) {
          int e, iz, ix;
          if (u < (S - 1) * K * (n & 1)) {
            e = u / K;
            iz = u - e * K;
            ix = n - 1;
          } else {
            e = S - 1;
            iz = u - (S - 1) * K * (n & 1);
            ix = K - 1;
          }
          goff[jg] = rb[1] + ((unsigned)(jx0 + e) * n3 + (unsigned)(iz * n2 + K * n + ix)) * 8;
          gq[jg] = (unsigned)(SH::rY + e * (K * n) + iz * n + ix);
        } else if ((u -= NYS) < NWS) {
          const bool last = u >= (S - 1) * (n & 1);
          const int e = last ? S - 1 : u, ix = last ? K - 1 : n - 1;
          goff[jg] = rb[3] + ((unsigned)(jx0 + e) * n3 + (unsigned)(K * n2 + K * n + ix)) * 8;
          gq[jg] = (unsigned)(SH::rW + e * n + ix);
        } else {  // the seam: ix = n-1 of element jx0-1 in the four rows
          const int q = u - NWS;
          const unsigned sb = (unsigned)jxs * n3 * 8;
          unsigned o;
          if (q < K * K) o = rb[0] + sb + (unsigned)((q / K) * n2 + (q % K) * n + K) * 8;
          else if (q < K * K + K) o = rb[1] + sb + (unsigned)((q - K * K) * n2 + K * n + K) * 8;
          else if (q < K * K + 2 * K) o = rb[2] + sb + (unsigned)(K * n2 + (q - K * K - K) * n + K) * 8;
          else o = rb[3] + sb + (unsigned)(K * n2 + K * n + K) * 8;
          goff[jg] = o;
          gq[jg] = (unsigned)(SH::rS + q);
        }
      }
      wg[jg] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, goff[jg], 0, 0));
    }
    if constexpr (LW == 2) {
#pragma unroll
      for (int j = JO; j < J; j++)
        v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off[j], 0, 0));
    }
  } else if constexpr (LW == 1) {
    const unsigned row_d = rb[0] / 8;  // row base in doubles
#pragma unroll
    for (int jw = 0; jw < JW; jw++) {
      const int t = jw * kStreamThreads + tid;
      unsigned o = kOOB, q = 0;
      if (t < S * PE) {
        const int e = t / PE, i = t - (t / PE) * PE;
        const unsigned eb = (unsigned)(jx0 + e) * n3;  // element base within the row (doubles)
        const unsigned par = (row_d + eb) & 1u;
        if (par || i < PE - 1) {
          o = rb[0] + (eb + 2u * i - par) * 8;
          q = (unsigned)(e * L + 2 * i + 1) - par;  // >= 0: the high half lies in the run or just past it
          if (!(par && i == 0)) q |= 1u << 16;
          if (!(par && i == PE - 1)) q |= 1u << 17;
        }
      }
      wq[jw] = q;
      w[jw] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, o, 0, 0);
    }
#pragma unroll
    for (int j = JO; j < J; j++)
      v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off[j], 0, 0));
  } else {
#pragma unroll
    for (int j = 0; j < J; j++)
      v[j] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off[j], 0, 0));
  }

  // ---- shared-node table (computed while the loads are in flight): the 8
  // LDS positions of each shared node this thread sums (16 bits each)
  unsigned np[NN][4];
  unsigned nflag[NN];  // bit0 cx, bit1 cy, bit2 cz, bit3 active
#pragma unroll
  for (int i = 0; i < NN; i++) {
    const int kk = i * kStreamThreads + tid;
    nflag[i] = 0;
    np[i][0] = np[i][1] = np[i][2] = np[i][3] = 0;
    if (kk >= S * SH::nshared) continue;
    const int e = kk / SH::nshared;
    int r = kk - e * SH::nshared;
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
  // MODE 1, 2: axes on which this item holds the rank surface (its partners
  // across it are the wrap partners: segment 0, row jy = 0, plane jz = 0)
  const unsigned surf = MODE ? ((jx0 == 0 ? 1u : 0u) | (jy == 0 ? 2u : 0u) | (jz == 0 ? 4u : 0u)) : 0u;

  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  if constexpr (LW >= 2) {
#pragma unroll
    for (int jr = 0; jr < JR; jr++) {
      if (roff[jr] == kOOB) continue;
      const u32x2 lo = {wr[jr][0], wr[jr][1]}, hi = {wr[jr][2], wr[jr][3]};
      lds[rq[jr]] = __builtin_bit_cast(double, lo);
      lds[rq[jr] + 1] = __builtin_bit_cast(double, hi);
    }
#pragma unroll
    for (int jg = 0; jg < JG; jg++)
      if (goff[jg] != kOOB) lds[gq[jg]] = wg[jg];
    if constexpr (LW == 2) {
#pragma unroll
      for (int j = JO; j < J; j++) lds[j * kStreamThreads + tid] = v[j];
    }
  } else if constexpr (LW == 1) {
#pragma unroll
    for (int jw = 0; jw < JW; jw++) {
      const unsigned q = wq[jw] & 0xffffu;
      const u32x2 lo = {w[jw][0], w[jw][1]}, hi = {w[jw][2], w[jw][3]};
      if (wq[jw] & (1u << 16)) lds[q - 1] = __builtin_bit_cast(double, lo);
      if (wq[jw] & (1u << 17)) lds[q] = __builtin_bit_cast(double, hi);
    }
#pragma unroll
    for (int j = JO; j < J; j++) lds[j * kStreamThreads + tid] = v[j];
  } else {
#pragma unroll
    for (int j = 0; j < J; j++) lds[j * kStreamThreads + tid] = v[j];
  }
  // MODE 1: one bit per LDS position, set for the copies of the rank-surface
  // nodes, which the item leaves to the surface sums and does not store
  // either: the surface sums run at the same time (the surface blocks)
  unsigned *const smark = reinterpret_cast<unsigned *>(lds + SH::P);
  if (MODE == 1 && surf) {
    for (int w = tid; w < kFlowMarkWords<N, S>; w += kStreamThreads) smark[w] = 0u;
  }
  __syncthreads();
  // ---- shared nodes: x pairs, then y, then z (TREE 1: z, y, x); the sum
  // goes to every copy
#pragma unroll
  for (int i = 0; i < NN; i++) {
    const unsigned f = nflag[i];
    if (!(f & 8u)) continue;
    // a node whose x pair (element 0 of the item only), y or z pair crosses
    // the rank surface is the surface sums'
    if (MODE && (surf & ((f & 6u) | ((f & 17u) == 17u ? 1u : 0u)))) {
      if constexpr (MODE == 1) {
        const unsigned fc = f & 7u;
#pragma unroll
        for (int c = 0; c < 8; c++) {
          if ((c & ~fc) != 0u) continue;  // copies the node has
          const unsigned pc = (c & 1) ? np[i][c >> 1] >> 16 : np[i][c >> 1] & 0xffff;
          atomicOr(&smark[pc >> 5], 1u << (pc & 31));
        }
      }
      continue;
    }
    // copy c = sx | sy << 1 | sz << 2 at position p<c>; tree 1 exchanges sx and sz
    const unsigned p0 = np[i][0] & 0xffff, p1 = TREE ? np[i][2] & 0xffff : np[i][0] >> 16;
    const unsigned p2 = np[i][1] & 0xffff, p3 = TREE ? np[i][3] & 0xffff : np[i][1] >> 16;
    const unsigned p4 = TREE ? np[i][0] >> 16 : np[i][2] & 0xffff, p5 = np[i][2] >> 16;
    const unsigned p6 = TREE ? np[i][1] >> 16 : np[i][3] & 0xffff, p7 = np[i][3] >> 16;
    const bool cx = f & (TREE ? 4u : 1u), cy = f & 2u, cz = f & (TREE ? 1u : 4u);
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
  // MODE 2: the rank-surface nodes of this item (block-uniform: only items on
  // the surface have any); each side of the surface sums its copies, and
  // those across the surface where the rank has a neighbour on that side
  if constexpr (MODE == 2) {
    if (surf) {
#pragma unroll
      for (int i = 0; i < NN; i++) {
        const unsigned f = nflag[i];
        if (!(f & 8u)) continue;
        const unsigned wrap = surf & ((f & 6u) | ((f & 17u) == 17u ? 1u : 0u));
        if (!wrap) continue;
        // copy (sx | sy << 1 | sz << 2), side 1 = the partner across that axis;
        // all eight are read (positions of copies the node does not have lie
        // in the footprint too, and go unused).  Tree 1: the axes x and z
        // exchanged in the copy numbering, the flags, the split axes and the
        // neighbour mask, so that the levels below run z, y, x.
        const unsigned pp[8] = {
            np[i][0] & 0xffff, TREE ? np[i][2] & 0xffff : np[i][0] >> 16,
            np[i][1] & 0xffff, TREE ? np[i][3] & 0xffff : np[i][1] >> 16,
            TREE ? np[i][0] >> 16 : np[i][2] & 0xffff, np[i][2] >> 16,
            TREE ? np[i][1] >> 16 : np[i][3] & 0xffff, np[i][3] >> 16};
        const unsigned ft = TREE ? (f & 2u) | ((f & 1u) << 2) | ((f >> 2) & 1u) : f;
        const unsigned wt = TREE ? (wrap & 2u) | ((wrap & 1u) << 2) | ((wrap >> 2) & 1u) : wrap;
        const int nbr = TREE ? (I.nbr_mask & 12) | ((I.nbr_mask & 3) << 4) | ((I.nbr_mask >> 4) & 3) : I.nbr_mask;
        double c[8];
#pragma unroll
        for (int m = 0; m < 8; m++) c[m] = lds[pp[m]];
        // pair[a][side]: that side of axis a sums both copies -- a pair inside
        // the rank, or across the surface with a neighbour on that side
        bool pair[3][2];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int sd = 0; sd < 2; sd++)
            pair[a][sd] = (ft & (1u << a)) && (!(wt & (1u << a)) || (nbr & (1 << (2 * a + sd))));
        // the tree level by level for both sides of every axis
        double A[2][2][2], Q[2][2][2], R[2][2][2];
#pragma unroll
        for (int gx = 0; gx < 2; gx++)
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const double x0 = c[2 * k], x1 = c[2 * k + 1];
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
          const bool ex = (!(m & 1) || (ft & 1u)) && (!(m & 2) || (ft & 2u)) && (!(m & 4) || (ft & 4u));
          if (!ex) continue;
          // copy m's side on each split axis
          const bool gx = (wt & 1u) && (m & 1), gy = (wt & 2u) && (m & 2), gz = (wt & 4u) && (m & 4);
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
  // MODE 3: the rank-surface nodes of this item, each side of the surface a
  // node of its own that sums its local copies and the received ones
  if constexpr (MODE == 3) {
    if (surf) {
#pragma unroll
      for (int i = 0; i < NN; i++) {
        const unsigned f = nflag[i];
        if (!(f & 8u)) continue;
        const unsigned wrap = surf & ((f & 6u) | ((f & 17u) == 17u ? 1u : 0u));
        if (!wrap) continue;
        // the node's element and local indices (the node table's enumeration)
        const int kk = i * kStreamThreads + tid;
        const int e = kk / SH::nshared;
        int r = kk - e * SH::nshared;
        int ix, iy, iz;
        if (r < K * K) {
          iz = 0; iy = r / K; ix = r - (r / K) * K;
        } else if (r < K * K + K * (K - 1)) {
          r -= K * K;
          iz = 1 + r / K; iy = 0; ix = r - (r / K) * K;
        } else {
          r -= K * K + K * (K - 1);
          constexpr int K1 = (K > 1) ? K - 1 : 1;
          iz = 1 + r / K1; iy = 1 + r - (r / K1) * K1; ix = 0;
        }
        const unsigned pp[8] = {np[i][0] & 0xffff, np[i][0] >> 16, np[i][1] & 0xffff, np[i][1] >> 16,
                                np[i][2] & 0xffff, np[i][2] >> 16, np[i][3] & 0xffff, np[i][3] >> 16};
        const unsigned fc = f & 7u;
        // own copy and partner copy (bit set) per axis: local index, element
        const int jx = jx0 + e;
        const int oi[3] = {ix, iy, iz}, oj[3] = {jx, jy, jz};
        const int pj[3] = {e > 0 ? jx - 1 : jxs, jym, jzm};
        const int m3[3] = {G.mx, G.my, G.mz};
#pragma unroll
        for (int sd = 0; sd < 8; sd++) {
          if (sd & ~(int)wrap) continue;  // one node per side of each split axis
          // pairs this side sums: inside the rank, or across the surface
          // with a neighbour on that side
          bool pa[3];
#pragma unroll
          for (int a = 0; a < 3; a++)
            pa[a] = (fc & (1u << a)) && (!(wrap & (1u << a)) || (I.nbr_mask & (1 << (2 * a + ((sd >> a) & 1)))));
          // leaves: bit a of l set = the partner copy on axis a (the received
          // copy on a split axis)
          double lv[8];
#pragma unroll
          for (int l = 0; l < 8; l++) {
            lv[l] = 0.0;
            if ((l & 1) && !pa[0]) continue;
            if ((l & 2) && !pa[1]) continue;
            if ((l & 4) && !pa[2]) continue;
            const int dm = l & (int)wrap;  // received on these axes
            if (!dm) {
              lv[l] = lds[pp[l | sd]];
              continue;
            }
            int ci[3], cj[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
              if (wrap & (1u << a)) {  // the side's own surface copy
                const bool hi = (sd >> a) & 1;
                ci[a] = hi ? K : 0;
                cj[a] = hi ? m3[a] - 1 : 0;
              } else {
                const bool pt = (l >> a) & 1;
                ci[a] = pt ? K : oi[a];
                cj[a] = pt ? pj[a] : oj[a];
              }
            }
            lv[l] = W.recv[recv_index(W.sl, n, G.mx, G.my, dm & 1, (dm >> 1) & 1, (dm >> 2) & 1, sd & 1,
                                      (sd >> 1) & 1, (sd >> 2) & 1, ci[0], ci[1], ci[2], cj[0], cj[1], cj[2])];
          }
          // x pairs, then y, then z
          double q[2];
#pragma unroll
          for (int lz = 0; lz < 2; lz++) {
            const double p0 = pa[0] ? lv[4 * lz] + lv[4 * lz + 1] : lv[4 * lz];
            const double p1 = pa[0] ? lv[4 * lz + 2] + lv[4 * lz + 3] : lv[4 * lz + 2];
            q[lz] = pa[1] ? p0 + p1 : p0;
          }
          const double res = pa[2] ? q[0] + q[1] : q[0];
          // every local copy of this side
#pragma unroll
          for (int nw = 0; nw < 8; nw++) {
            if (nw & (int)wrap) continue;
            if (nw & ~(int)fc) continue;
            lds[pp[nw | sd]] = res;
          }
        }
      }
    }
  }
  __syncthreads();
  // ---- stores of the locations this item owns; the others go out of range
  // (MODE 1: the copies of rank-surface nodes too)
  const bool excl = MODE == 1 && surf;
  auto marked = [&](unsigned pos) { return ((smark[pos >> 5] >> (pos & 31)) & 1u) != 0u; };
  // the 8-byte slots stored below: OWN and partner regions (LW 0, 1), the
  // partner regions (LW 2), none (LW 3: every owned location is in a pair or
  // single slot)
  constexpr int J8 = LW == 3 ? J : LW == 2 ? JO : 0;
  if constexpr (LW >= 2) {
    if (excl) {
      // (block-uniform) a pair with a rank-surface half stores its other half alone
#pragma unroll
      for (int jr = 0; jr < JR; jr++) {
        const u32x2 lo = __builtin_bit_cast(u32x2, lds[rq[jr]]), hi = __builtin_bit_cast(u32x2, lds[rq[jr] + 1]);
        const bool used = roff[jr] != kOOB;
        const bool m0 = used && marked(rq[jr]), m1 = used && marked(rq[jr] + 1);
        if (m0 || m1) {
          if (!m0) __builtin_amdgcn_raw_buffer_store_b64(lo, rsrc, roff[jr], 0, 0);
          if (!m1) __builtin_amdgcn_raw_buffer_store_b64(hi, rsrc, roff[jr] + 8u, 0, 0);
        } else {
          const u32x4 x = {lo[0], lo[1], hi[0], hi[1]};
          __builtin_amdgcn_raw_buffer_store_b128(x, rsrc, roff[jr], 0, 0);
        }
      }
#pragma unroll
      for (int jg = 0; jg < JG; jg++) {
        const unsigned o = (goff[jg] != kOOB && marked(gq[jg])) ? kOOB : goff[jg];
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, lds[gq[jg]]), rsrc, o, 0, 0);
      }
    } else {
#pragma unroll
      for (int jr = 0; jr < JR; jr++) {
        const u32x2 lo = __builtin_bit_cast(u32x2, lds[rq[jr]]), hi = __builtin_bit_cast(u32x2, lds[rq[jr] + 1]);
        const u32x4 x = {lo[0], lo[1], hi[0], hi[1]};
        __builtin_amdgcn_raw_buffer_store_b128(x, rsrc, roff[jr], 0, 0);
      }
#pragma unroll
      for (int jg = 0; jg < JG; jg++)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, lds[gq[jg]]), rsrc, goff[jg], 0, 0);
    }
  }
  if (excl) {
#pragma unroll
    for (int j = J8; j < J; j++)
      if (marked(j * kStreamThreads + tid)) keep &= ~(1u << j);
  }
#pragma unroll
  for (int j = J8; j < J; j++) {
    const u32x2 x = __builtin_bit_cast(u32x2, lds[j * kStreamThreads + tid]);
    __builtin_amdgcn_raw_buffer_store_b64(x, rsrc, (keep & (1u << j)) ? off[j] : kOOB, 0, 0);
  }
}

// Launches k_flow<n, S, lw, mode, tree> on nb item blocks of kStreamThreads
// (faces_flow.hip), and for mode 1 the surface blocks W->nblk after them (W
// may be null: none); mode 3 reads its receive slots from W.  Tree 1 is
// compiled with load widths 2 and 3 and modes 0 and 2, mode 3 with load
// width 2.
int launch_flow(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                const SurfWork *W, hipStream_t st);

}  // namespace faces
