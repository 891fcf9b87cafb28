// faces_flow_g0.hip -- the one-item kernel's launches for n = 2, 3
// (faces_flow_launch.hpp), compiled apart from the other element sizes.
#include "faces_flow_launch.hpp"

namespace faces {

FACES_FLOW_GROUP(launch_flow_g0, FACES_FLOW_CASE(2) FACES_FLOW_CASE(3))

}  // namespace faces
