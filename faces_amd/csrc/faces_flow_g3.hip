// faces_flow_g3.hip -- the one-item kernel's launches for n = 9, 10, 11, 12
// (faces_flow_launch.hpp), compiled apart from the other element sizes.
#include "faces_flow_launch.hpp"

namespace faces {

FACES_FLOW_GROUP(launch_flow_g3, FACES_FLOW_CASE(9) FACES_FLOW_CASE(10) FACES_FLOW_CASE(11) FACES_FLOW_CASE(12))

}  // namespace faces
