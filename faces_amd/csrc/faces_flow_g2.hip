// faces_flow_g2.hip -- the one-item kernel's launches for n = 6, 7, 8
// (faces_flow_launch.hpp), compiled apart from the other element sizes.
#include "faces_flow_launch.hpp"

namespace faces {

FACES_FLOW_GROUP(launch_flow_g2, FACES_FLOW_CASE(6) FACES_FLOW_CASE(7) FACES_FLOW_CASE(8))

}  // namespace faces
