// Toeplitz kernel construction from a trajectory (placeholder until the
// spreading kernels land).
#include "internal.hpp"

namespace msr {
void op_build_from_traj(msr_normalop*, const msr_traj*, const void*, int, hipStream_t) {
    fail(MSR_ERR_UNSUPPORTED, "kernel construction from a trajectory is not built yet");
}
}  // namespace msr
