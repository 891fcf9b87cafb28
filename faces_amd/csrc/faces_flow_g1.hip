// faces_flow_g1.hip -- the one-item kernel's launches for n = 4, 5
// (faces_flow_launch.hpp), compiled apart from the other element sizes.
#include "faces_flow_launch.hpp"

namespace faces {

FACES_FLOW_GROUP(launch_flow_g1, FACES_FLOW_CASE(4) FACES_FLOW_CASE(5))

}  // namespace faces
