// Pass C (3D) instantiations for K = 6.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(6)
}  // namespace msr
