// Pass B instantiations for K = 8 (split by K so the build runs in parallel).
#include "apply2d_b.hpp"

namespace msr {
MSR_DEFINE_PASSB(8)
}  // namespace msr
