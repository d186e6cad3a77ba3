// Subspace backprojection (placeholder until the spreading kernels land).
#include "internal.hpp"

namespace msr {
void backprojection(msr_ctx*, const msr_traj*, int, const int64_t*, const void*, int, int,
                    const void*, int, const void* const*, int, const float*, void*, hipStream_t) {
    fail(MSR_ERR_UNSUPPORTED, "backprojection is not built yet");
}
}  // namespace msr
