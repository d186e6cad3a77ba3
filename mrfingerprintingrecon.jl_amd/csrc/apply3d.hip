// Fused 3D Toeplitz subspace normal operator.
#include "apply3d.hpp"

namespace msr {

void alloc_scratch_3d(msr_normalop*) {
    fail(MSR_ERR_UNSUPPORTED, "3D Toeplitz apply is not built yet");
}
void launch_apply3d_passes(msr_normalop*, const float2*, hipStream_t, const int*) {
    fail(MSR_ERR_UNSUPPORTED, "3D Toeplitz apply is not built yet");
}
void launch_apply3d_passes_coil(msr_normalop*, const float2*, int, hipStream_t) {
    fail(MSR_ERR_UNSUPPORTED, "3D Toeplitz apply is not built yet");
}

}  // namespace msr
