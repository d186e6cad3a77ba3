// Fused 3D Toeplitz subspace normal operator (five passes over the 2N grid).
#pragma once

#include "internal.hpp"

namespace msr {

// b, T1, T2 and partial in ComplexF64 when f64, else ComplexF32
void launch_apply3d_passes(msr_normalop* op, const void* b, bool f64, hipStream_t st,
                           const int* skip);
void launch_apply3d_passes_coil(msr_normalop* op, const float2* b, int coil, hipStream_t st);

}  // namespace msr
