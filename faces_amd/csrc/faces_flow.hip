// faces_flow.hip -- launches of the one-item kernel (faces_flow.hpp).
//
// The plan (faces_plan.hip) chooses S, the load width and the grid;
// launch_flow maps them to the compiled kernel.  The instantiations of each
// group of element sizes live in a translation unit of their own
// (faces_flow_g*.hip), so that they compile beside one another.
#include "faces_flow_launch.hpp"

namespace faces {

int launch_flow(int n, int S, int lw, int mode, int tree, const Geom &G, const FlowItems &F, unsigned nb,
                const SurfWork *W, hipStream_t st)
{
  if (n == 2 || n == 3) return launch_flow_g0(n, S, lw, mode, tree, G, F, nb, W, st);
  if (n == 4 || n == 5) return launch_flow_g1(n, S, lw, mode, tree, G, F, nb, W, st);
  if (n >= 6 && n <= 8) return launch_flow_g2(n, S, lw, mode, tree, G, F, nb, W, st);
  if (n >= 9 && n <= 12) return launch_flow_g3(n, S, lw, mode, tree, G, F, nb, W, st);
  return fail(FACES_ERR_ARG, "one-item kernel has no n=%d specialisation", n);
}

}  // namespace faces
