// Pass C (3D) instantiations for K = 4.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(4)
}  // namespace msr
