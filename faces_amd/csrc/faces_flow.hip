// faces_flow.hip -- launches of the one-item kernel (faces_flow.hpp).
//
// Its instantiations (element size n, elements per item S, load width, mode,
// tree order) live in this translation unit of their own, so that they
// compile beside the plan's.  The plan (faces_plan.hip) chooses S, the load
// width and the grid; launch_flow maps them to the compiled kernel.
#include "faces_common.hpp"
#include "faces_flow.hpp"

namespace faces {

namespace {

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
    // alignment), periodic or open
    if (lw != 2 || mode == 1)
      return fail(FACES_ERR_ARG, "one-item kernel: tree order z,y,x needs load width 2 and mode 0 or 2");
    if (mode == 2) FACES_LAUNCH_FLOW(2, 2, 1);
    else FACES_LAUNCH_FLOW(2, 0, 1);
  } else if (mode == 3) {
    FACES_LAUNCH_FLOW(2, 3, 0);
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
  if constexpr (N == 2) {
    if (S == 32) return launch_flow_ns<N, 32>(lw, mode, tree, G, F, nb, W, st);
  }
  return fail(FACES_ERR_ARG, "one-item kernel has no n=%d S=%d specialisation", N, S);
}

}  // namespace

int launch_flow(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                const SurfWork *W, hipStream_t st)
{
  switch (n) {
    case 2: return launch_flow_n<2>(S, lw, mode, tree, G, F, nb, W, st);
    case 3: return launch_flow_n<3>(S, lw, mode, tree, G, F, nb, W, st);
    case 4: return launch_flow_n<4>(S, lw, mode, tree, G, F, nb, W, st);
    case 5: return launch_flow_n<5>(S, lw, mode, tree, G, F, nb, W, st);
    case 6: return launch_flow_n<6>(S, lw, mode, tree, G, F, nb, W, st);
    case 7: return launch_flow_n<7>(S, lw, mode, tree, G, F, nb, W, st);
    case 8: return launch_flow_n<8>(S, lw, mode, tree, G, F, nb, W, st);
    case 9: return launch_flow_n<9>(S, lw, mode, tree, G, F, nb, W, st);
    case 10: return launch_flow_n<10>(S, lw, mode, tree, G, F, nb, W, st);
    case 11: return launch_flow_n<11>(S, lw, mode, tree, G, F, nb, W, st);
    case 12: return launch_flow_n<12>(S, lw, mode, tree, G, F, nb, W, st);
    default: return fail(FACES_ERR_ARG, "one-item kernel has no n=%d specialisation", n);
  }
}

}  // namespace faces
