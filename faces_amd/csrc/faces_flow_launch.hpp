// faces_flow_launch.hpp -- the launches of the one-item kernel (faces_flow.hpp)
// for one element size n: the kernel instantiations (elements per item S,
// load width, mode, tree order) of each n are compiled in the translation
// unit of its group of sizes (faces_flow_n*.hip), beside one another.
#pragma once

#include "faces_common.hpp"
#include "faces_flow.hpp"

namespace faces {

template <int N, int S>
int launch_flow_ns(int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                   const SurfWork *W, hipStream_t st)
{
  // the footprint, and for mode 1 the marks of the rank-surface copies
  const size_t lds = (size_t)StreamShape<N, S>::P * sizeof(double) +
                     (mode == 1 ? (size_t)kFlowMarkWords<N, S> * sizeof(unsigned) : 0);
  SurfWork sw = {};
  unsigned nblk = 0;
  if (mode == 3 && (!W || W->nblk || lw != 2 || tree))
    return fail(FACES_ERR_ARG, "one-item kernel: mode 3 takes the receive slots, load width 2, no surface blocks");
  if (W) {
    if (mode != 1 && mode != 3) return fail(FACES_ERR_ARG, "one-item kernel: surface blocks need mode 1");
    sw = *W;
    // leading surface blocks: a multiple of 8, so that item block b still
    // runs on XCD b % 8
    nblk = W->lead ? (W->nblk + 7u) & ~7u : W->nblk;
    sw.nblk = nblk;
  }
  sw.sfirst = W && W->lead ? 0u : nb;
  sw.item0 = W && W->lead ? nblk : 0u;
  if ((unsigned long)nb + nblk > 0x7fffffffUL) return fail(FACES_ERR_ARG, "one-item kernel: grid too large");
  const dim3 grid(nb + nblk), blk(kStreamThreads);
#define FACES_LAUNCH_FLOW(LW, MODE, TREE) k_flow<N, S, LW, MODE, TREE><<<grid, blk, lds, st>>>(G, F, sw)
  if (tree) {
    // the 6-message schedule's tree: owned runs in 16-byte pairs (any
    // alignment; load width 3: the partner regions' too), periodic or open
    if (lw < 2 || mode == 1 || mode == 3)
      return fail(FACES_ERR_ARG, "one-item kernel: tree order z,y,x needs load width 2 or 3 and mode 0 or 2");
    if (lw == 3) {
      if (mode == 2) FACES_LAUNCH_FLOW(3, 2, 1);
      else FACES_LAUNCH_FLOW(3, 0, 1);
    } else if (mode == 2) {
      FACES_LAUNCH_FLOW(2, 2, 1);
    } else {
      FACES_LAUNCH_FLOW(2, 0, 1);
    }
  } else if (mode == 3) {
    FACES_LAUNCH_FLOW(2, 3, 0);
  } else if (lw == 3) {
    // the partner regions in 16-byte pairs too
    if (mode == 2) FACES_LAUNCH_FLOW(3, 2, 0);
    else if (mode == 1) FACES_LAUNCH_FLOW(3, 1, 0);
    else FACES_LAUNCH_FLOW(3, 0, 0);
  } else if (mode == 2) {
    if (lw == 2) FACES_LAUNCH_FLOW(2, 2, 0);
    else if (lw == 1) FACES_LAUNCH_FLOW(1, 2, 0);
    else FACES_LAUNCH_FLOW(0, 2, 0);
  } else if (mode == 1) {
    if (lw == 2) FACES_LAUNCH_FLOW(2, 1, 0);
    else if (lw == 1) FACES_LAUNCH_FLOW(1, 1, 0);
    else FACES_LAUNCH_FLOW(0, 1, 0);
  } else {
    if (lw == 2) FACES_LAUNCH_FLOW(2, 0, 0);
    else if (lw == 1) FACES_LAUNCH_FLOW(1, 0, 0);
    else FACES_LAUNCH_FLOW(0, 0, 0);
  }
#undef FACES_LAUNCH_FLOW
  FACES_HIP(hipGetLastError());
  return FACES_OK;
}

template <int N>
int launch_flow_n(int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                  const SurfWork *W, hipStream_t st)
{
  switch (S) {
    case 1: return launch_flow_ns<N, 1>(lw, mode, tree, G, F, nb, W, st);
    case 2: return launch_flow_ns<N, 2>(lw, mode, tree, G, F, nb, W, st);
    case 3: return launch_flow_ns<N, 3>(lw, mode, tree, G, F, nb, W, st);
    case 4: return launch_flow_ns<N, 4>(lw, mode, tree, G, F, nb, W, st);
    default: break;
  }
  if constexpr (N >= 5 && N <= 8) {
    if (S == 5) return launch_flow_ns<N, 5>(lw, mode, tree, G, F, nb, W, st);
  }
  if constexpr (N == 5) {
    if (S == 10) return launch_flow_ns<N, 10>(lw, mode, tree, G, F, nb, W, st);
  }
  if constexpr (N <= 7) {
    if (S == 6) return launch_flow_ns<N, 6>(lw, mode, tree, G, F, nb, W, st);
  }
  if constexpr (N <= 6) {
    if (S == 8) return launch_flow_ns<N, 8>(lw, mode, tree, G, F, nb, W, st);
  }
  if constexpr (N <= 4) {
    if (S == 16) return launch_flow_ns<N, 16>(lw, mode, tree, G, F, nb, W, st);
  }
  if constexpr (N <= 3) {
    if (S == 32) return launch_flow_ns<N, 32>(lw, mode, tree, G, F, nb, W, st);
  }
  if constexpr (N == 2) {
    if (S == 64) return launch_flow_ns<N, 64>(lw, mode, tree, G, F, nb, W, st);
  }
  return fail(FACES_ERR_ARG, "one-item kernel has no n=%d S=%d specialisation", N, S);
}


// The n of one group, dispatched to launch_flow_n<N>.
#define FACES_FLOW_GROUP(NAME, ...)                                                                       \
  int NAME(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb, \
           const SurfWork *W, hipStream_t st)                                                          \
  {                                                                                                    \
    switch (n) {                                                                                       \
      __VA_ARGS__                                                                                      \
      default: return fail(FACES_ERR_ARG, "one-item kernel has no n=%d specialisation", n);       \
    }                                                                                                  \
  }
#define FACES_FLOW_CASE(NN) \
  case NN: return launch_flow_n<NN>(S, lw, mode, tree, G, F, nb, W, st);

int launch_flow_g0(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                   const SurfWork *W, hipStream_t st);  // n = 2, 3
int launch_flow_g1(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                   const SurfWork *W, hipStream_t st);  // n = 4, 5
int launch_flow_g2(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                   const SurfWork *W, hipStream_t st);  // n = 6, 7, 8
int launch_flow_g3(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                   const SurfWork *W, hipStream_t st);  // n = 9 to 12

}  // namespace faces
