// Fused 3D Toeplitz subspace normal operator (five passes over the 2N grid).
#pragma once

#include "internal.hpp"

namespace msr {

void alloc_scratch_3d(msr_normalop* op);
void launch_apply3d_passes(msr_normalop* op, const float2* b, hipStream_t st, const int* skip);
void launch_apply3d_passes_coil(msr_normalop* op, const float2* b, int coil, hipStream_t st);

}  // namespace msr
