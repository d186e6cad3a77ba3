// Pass C (3D) instantiations for K = 7.
#include "apply3d_c.hpp"

namespace msr {
MSR_DEFINE_PASSC3(7)
}  // namespace msr
