// Pass C (3D) instantiations for K = 2.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(2)
}  // namespace msr
