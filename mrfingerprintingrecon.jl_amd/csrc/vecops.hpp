// Vector kernels: apply epilogue (group reduction with alpha/beta), and the
// device conjugate-gradient steps with FP64 reductions.
//
// Each CG step is a per-operator block body taking (block, number of blocks)
// of its operator, so that the same arithmetic runs in a single-operator
// launch and in a batched launch that concatenates the blocks of many
// operators (each block finds its operator in a job table): every operator's
// reductions see the same block partition, hence the same rounding, alone and
// in a batch.
#pragma once

#include "internal.hpp"

namespace msr {

constexpr int RED_BLOCKS = 512;   // blocks of the two-level FP64 reductions
constexpr int RED_THREADS = 256;
// reduction partials of a single-operator CG: the dot partials (two per block)
// in [0, 2 RED_BLOCKS), the residual partials of the fused step in
// [2 RED_BLOCKS, 3 RED_BLOCKS)
constexpr int CG_PART_DOUBLES = 3 * RED_BLOCKS;

inline int red_grid(int64_t n) {
    int64_t g = (n + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
    if (g < 1) g = 1;
    if (g > RED_BLOCKS) g = RED_BLOCKS;
    return (int)g;
}

// x = beta*x + alpha*scale*sum_g partial[g]   (beta_zero: overwrite); CT is
// the complex type of partial and x (the apply's arithmetic type)
template <class CT>
void launch_epilogue(const CT* partial, int G, int64_t n, CT* x, CT alpha, CT beta, bool beta_zero,
                     double scale, hipStream_t st, const int* skip);

// x = beta*x (or 0 when beta_zero)
template <class CT>
void launch_scale(CT* x, int64_t n, CT beta, bool beta_zero, hipStream_t st);

// CG building blocks (scalars in CgScalars on the device).  VT is the type of
// the solution / residual / search vectors: float2 (ComplexF32 throughout) or
// double2 (FP64 vectors).  AT is the arithmetic type of the apply: float2
// (the apply reads an FP32 copy of u and returns c in FP32) or double2 (with
// VT = double2: the apply reads u itself and returns c in FP64).
template <class VT, class AT>
struct CgVecs {
    VT* x;          // solution
    VT* r;          // residual
    VT* u;          // search direction
    AT* ua;         // apply input: u itself when AT == VT, else an AT copy
    AT* c;          // A u
};

template <class VT, class AT>
void launch_cg_start(const float2* b, const float2* x_in, const CgVecs<VT, AT>& v, int64_t n,
                     bool zero_x, double* part, CgScalars* S, double reltol, double abstol,
                     int maxiter, bool r_minus_c, hipStream_t st);
template <class VT, class AT>
void launch_cg_u(const CgVecs<VT, AT>& v, int64_t n, CgScalars* S, hipStream_t st);
// c = scale*sum_g partial[g]; dot(u, c) in FP64; alpha = res^2 / dot
// (step_kernel = false: no alpha kernel; launch_cg_xr_step forms alpha)
template <class VT, class AT>
void launch_cg_reduce_dot(const AT* partial, int G, int64_t n, const CgVecs<VT, AT>& v,
                          double scale, double* part, CgScalars* S, hipStream_t st,
                          bool step_kernel = true);
template <class VT, class AT>
void launch_cg_xr(const CgVecs<VT, AT>& v, int64_t n, double* part, CgScalars* S, hipStream_t st);
// x += alpha u; r -= alpha c with alpha = res^2 / dot formed by every block
// from the dot partials in part (the same value in each: one wavefront sums
// them in a fixed order), so no alpha kernel runs; the residual partials go to
// part + 2 RED_BLOCKS (part holds CG_PART_DOUBLES)
template <class VT, class AT>
void launch_cg_xr_step(const CgVecs<VT, AT>& v, int64_t n, double* part, CgScalars* S,
                       hipStream_t st);
// x32 = (float) x64
void launch_to_f32(const double2* x64, float2* x32, int64_t n, hipStream_t st);

// ---------------------------------------------------------------- batched
// One operator's CG state in a batched launch.
template <class VT, class AT>
struct CgJob {
    CgVecs<VT, AT> v;
    const float2* b;         // right-hand side (start)
    const float2* x_in;      // initial guess (start, not initially zero)
    float2* x_out;           // FP32 result (finish, FP64 vectors)
    const AT* partial;       // apply group sums
    int G;                   // groups in partial
    double scale;            // apply normalisation 1/grid cells
    int64_t n;               // vector length V*K
    CgScalars* S;
    double* part;            // reduction partials (2 * RED_BLOCKS)
    AT* y;                   // target of the batched apply epilogue
    AT alpha, beta;          // batched apply epilogue
    int beta_zero;
};

// Block prefix of a batch: job j owns blocks [pre[j], pre[j+1]) of every
// batched vector launch, pre[j+1] - pre[j] = red_grid(n_j).
template <class VT, class AT>
void launch_cg_start_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                           bool zero_x, bool r_minus_c, double reltol, double abstol, int maxiter,
                           hipStream_t st);
template <class VT, class AT>
void launch_cg_u_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                       hipStream_t st);
template <class VT, class AT>
void launch_cg_reduce_dot_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                                hipStream_t st);
template <class VT, class AT>
void launch_cg_xr_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                        hipStream_t st);
template <class AT>
void launch_to_f32_batch(const CgJob<double2, AT>* jobs, const int* pre, int nj, int nblocks,
                         hipStream_t st);
// batched apply epilogue: y = beta y + alpha scale sum_g partial (per job)
template <class VT, class AT>
void launch_epilogue_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                           hipStream_t st);

}  // namespace msr
