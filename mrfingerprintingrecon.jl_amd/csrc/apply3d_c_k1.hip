// Pass C (3D) instantiations for K = 1.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(1)
}  // namespace msr
