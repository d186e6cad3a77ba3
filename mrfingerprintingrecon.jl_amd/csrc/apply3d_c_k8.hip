// Pass C (3D) instantiations for K = 8.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(8)
}  // namespace msr
