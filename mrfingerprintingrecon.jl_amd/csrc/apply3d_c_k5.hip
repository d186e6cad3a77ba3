// Pass C (3D) instantiations for K = 5.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(5)
}  // namespace msr
