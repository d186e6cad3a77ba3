// Fused 2D Toeplitz subspace normal operator (three passes over the 2N grid).
//
// Replaces the per-coil sequence of the reference GPU mul!
// (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:193-218: fill, pad, cuFFT,
// kernel_mul!, cuFFT, crop per coil) with three streaming passes in which
// every intermediate is written once and read once and the Toeplitz kernel is
// read once per apply for all coils:
//
//   A  (y-tile, k, c):  row y of S_c b_k, zero-padded to 2Nx, FFT along x;
//                       written transposed:  T[c][k][qx][y]   (y fastest)
//   B  (qx column, coil group): for each coil, the K columns T[c][:][qx][:]
//                       are zero-padded to 2Ny, FFT along y, multiplied by the
//                       per-voxel K x K Hermitian kernel (held in registers for
//                       the whole coil loop), inverse FFT along y, cropped to
//                       Ny and written back in place.
//   C  (y-tile, k, coil group): inverse FFT along x of rows of T, crop to Nx,
//                       multiply by conj(S_c), accumulate over the group's
//                       coils in registers; one partial image per group.
//   R  reduce:          x = beta x + alpha/(4 Nx Ny) sum_g partial_g.
#pragma once

#include "common.hpp"

namespace msr {

struct Apply2DArgs {
    int nx, ny;        // image size
    int K, C;          // coefficients, coils (C >= 1)
    const float2* cmaps;  // (C, ny, nx) or nullptr for the identity map
    const void* lam;      // dense kernel (2nx, P, 2ny), float2 or float
    int lam_real;
    float2* T;            // scratch (C, K, 2nx, ny)
    float2* partial;      // scratch (G, K, ny, nx)
    int G;                // coil groups of pass C (>= 1)
    int Gb;               // coil groups of pass B (>= 1)
    const float2* tw_x;   // twiddles for 2nx
    const float2* tw_y;   // twiddles for 2ny
};

// Launches passes A, B, C.  After this, partial[g] holds the group sums
// (unscaled).  The caller runs a reduction that applies alpha, beta and the
// 1/(4 nx ny) normalisation (see reduce kernels in cg.hpp / api).
void launch_apply2d_passes(const Apply2DArgs& a, const float2* b, hipStream_t st,
                           const int* skip_flag);

}  // namespace msr
