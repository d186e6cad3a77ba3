// Fused 3D Toeplitz subspace normal operator (five passes over the 2N grid).
#pragma once

#include "apply2d.hpp"
#include "internal.hpp"

namespace msr {

// am: ApplyMode (apply2d.hpp); b and partial in ComplexF64 unless AM_F32, T1
// and T2 in the mode's storage (cx64 for AM_F64C)
void launch_apply3d_passes(msr_normalop* op, const void* b, int am, hipStream_t st,
                           const int* skip);
void launch_apply3d_passes_coil(msr_normalop* op, const float2* b, int coil, hipStream_t st);

}  // namespace msr
