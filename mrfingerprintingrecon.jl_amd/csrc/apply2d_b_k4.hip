// Pass B instantiations for K = 4 (split by K so the build runs in parallel).
#include "apply2d_b.hpp"

namespace msr {
MSR_DEFINE_PASSB(4)
}  // namespace msr
