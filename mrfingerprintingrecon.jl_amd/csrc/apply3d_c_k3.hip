// Pass C (3D) instantiations for K = 3.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(3)
}  // namespace msr
