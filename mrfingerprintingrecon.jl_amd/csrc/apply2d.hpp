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
//
// Every pass is written as a per-operator block body that takes a linear block
// index local to its operator.  The same body serves two launch forms: one
// operator per launch, and a batched launch in which the blocks of several
// operators (same FFT lengths / K, any coil count) are concatenated and each
// block finds its operator in a job table.  An operator's arithmetic is
// therefore the same alone and inside a batch.
#pragma once

#include "common.hpp"

namespace msr {

// Apply arithmetic and intermediate storage: ComplexF32 throughout,
// ComplexF64 throughout, or ComplexF64 arithmetic with the FFT intermediates
// in the compact 64-bit storage (cx64, common.hpp).
enum ApplyMode { AM_F32 = 0, AM_F64 = 1, AM_F64C = 2 };

// Arithmetic is ComplexF32 (f64 = 0) or ComplexF64 (f64 = 1); the
// intermediate T has the arithmetic type, or cx64 when compact (with f64).
// The kernel and the coil maps are ComplexF32 in every case (promoted on
// load).  b, partial and the twiddles have the element type of the arithmetic.
struct Apply2DArgs {
    int nx, ny;           // image size
    int K, C;             // coefficients, coils (C >= 1)
    int f64;              // arithmetic precision
    int compact;          // T stored as cx64 (with f64)
    const void* b;        // input (K, ny, nx)
    const float2* cmaps;  // (C, ny, nx) or nullptr for the identity map
    const void* lam;      // dense kernel (2nx, P, 2ny), float2 or float
    int lam_real;
    void* T;              // scratch (C, K, 2nx, ny)
    void* partial;        // scratch (G, K, ny, nx)
    int G;                // coil groups of pass C (>= 1)
    int Gb;               // coil groups of pass B (>= 1)
    int order_a;          // pass A block order: 0 tile fastest, 1 coil fastest
    int order_c;          // pass C block order: 0 tile fastest, 1 coefficient fastest
    const void* tw_x;     // twiddles for 2nx
    const void* tw_y;     // twiddles for 2ny
    const int* skip;      // when non-null and *skip != 0 the passes do nothing
};

// rows per tile in passes A and C: 64-byte segments of the transposed
// intermediate for ComplexF32 (8 rows) and ComplexF64 (4 rows); with compact
// storage the rows are ComplexF64 in LDS, and 4 rows (32-byte segments, whose
// 128-byte lines neighbouring tiles on the same XCD complete in L2) keep three
// workgroups per CU, against one for 8 rows (measured 0.30 against 0.41 ms
// per apply at 256^2, 20 coils, K = 5)
#ifndef MSR_CX_ROWS
#define MSR_CX_ROWS 4
#endif
__host__ __device__ constexpr int tile_rows(int f64, int compact) {
    return !f64 ? 8 : (compact ? MSR_CX_ROWS : 4);
}
template <class CT, class ST>
constexpr int rows_per_tile() {
    return tile_rows(sizeof(CT) == 16, std::is_same<ST, cx64>::value);
}
__host__ __device__ inline int tiles2d(int ny, int rows) { return (ny + rows - 1) / rows; }
int pass_b_colblocks(int nx2, int ny2, int K, int f64);  // column blocks of pass B per coil group

// block counts per operator of each pass
inline int64_t blocks_a2d(const Apply2DArgs& a) {
    return (int64_t)tiles2d(a.ny, tile_rows(a.f64, a.compact)) * a.K * a.C;
}
inline int64_t blocks_c2d(const Apply2DArgs& a) {
    return (int64_t)tiles2d(a.ny, tile_rows(a.f64, a.compact)) * a.K * a.G;
}
inline int64_t blocks_b2d(const Apply2DArgs& a) {
    return (int64_t)pass_b_colblocks(2 * a.nx, 2 * a.ny, a.K, a.f64) * a.Gb;
}

// Launches passes A, B, C of one operator.  After this, partial[g] holds the
// group sums (unscaled); the caller's reduction applies alpha, beta and the
// 1/(4 nx ny) normalisation.
void launch_apply2d_passes(const Apply2DArgs& a, hipStream_t st);

// Batched form (am: ApplyMode).  The operators of one launch share the mode and 2nx
// (passes A, C) or (2ny, K, lam_real) (pass B); `jobs` is a device array of nj
// operators and `pre` a device array of nj + 1 block offsets (pre[0] = 0) for
// that pass.
void launch_pass_a_2d_batch(int nx2, int am, const Apply2DArgs* jobs, const int* pre, int nj,
                            int64_t nblocks, hipStream_t st);
void launch_pass_b_2d_batch(int ny2, int K, int lam_real, int am, const Apply2DArgs* jobs,
                            const int* pre, int nj, int64_t nblocks, hipStream_t st);
void launch_pass_c_2d_batch(int nx2, int am, const Apply2DArgs* jobs, const int* pre, int nj,
                            int64_t nblocks, hipStream_t st);

// Operator of block `blk` in a batched launch: the j with pre[j] <= blk < pre[j+1].
__device__ __forceinline__ int batch_job(const int* __restrict__ pre, int nj, int blk) {
    int lo = 0, hi = nj - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace msr
