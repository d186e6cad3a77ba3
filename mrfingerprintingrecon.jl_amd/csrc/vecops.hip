// Vector kernels and the device CG recurrence with FP64 reductions.
//
// The CG follows IterativeSolvers.jl 0.9 (called by the reference at
// test/reconstruct_radial_gpu.jl:99 and src/CoilMaps.jl:136):
//   u = r + (res^2/prev^2) u;  c = A u;  alpha = res^2 / dot(u, c);
//   x += alpha u;  r -= alpha c;  prev = res;  res = ||r||
// stopping when iter >= maxiter or res <= max(reltol*||r0||, abstol).
// Dot products, norms and step lengths are accumulated and kept in FP64 on the
// device, reduced in a fixed order (block partials, then one wavefront) so
// results are run-to-run deterministic.  The vectors are ComplexF32 or FP64
// (VT); the apply works in ComplexF32 or ComplexF64 (AT).
#include "apply2d.hpp"  // batch_job
#include "vecops.hpp"

namespace msr {

__device__ __forceinline__ double warp_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum over the block; result valid in thread 0.  s must hold blockDim/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* s) {
    v = warp_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) s[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += s[i];
    }
    return r;
}

// ------------------------------------------------------------------ epilogue
template <class CT>
__device__ __forceinline__ void epilogue_body(const CT* __restrict__ partial, int G, int64_t n,
                                              CT* __restrict__ x, CT alpha, CT beta, int beta_zero,
                                              double scale, int64_t blk, int64_t nblk) {
    using RT = real_t<CT>;
    for (int64_t i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
        CT s = partial[i];
        for (int g = 1; g < G; ++g) s = cadd(s, partial[(int64_t)g * n + i]);
        s = cscale(s, (RT)scale);
        CT v = cmul(alpha, s);
        if (!beta_zero) v = cadd(v, cmul(beta, x[i]));
        x[i] = v;
    }
}

template <class CT>
__global__ void epilogue_kernel(const CT* __restrict__ partial, int G, int64_t n,
                                CT* __restrict__ x, CT alpha, CT beta, int beta_zero,
                                double scale, const int* skip) {
    if (skip && *skip) return;
    epilogue_body(partial, G, n, x, alpha, beta, beta_zero, scale, blockIdx.x, gridDim.x);
}

template <class CT>
void launch_epilogue(const CT* partial, int G, int64_t n, CT* x, CT alpha, CT beta, bool beta_zero,
                     double scale, hipStream_t st, const int* skip) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(epilogue_kernel<CT>, dim3(grid), dim3(256), 0, st, partial, G, n, x, alpha,
                       beta, (int)beta_zero, scale, skip);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
__global__ void epilogue_batch_kernel(const CgJob<VT, AT>* __restrict__ jobs,
                                      const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT, AT>& J = jobs[j];
    epilogue_body(J.partial, J.G, J.n, J.y, J.alpha, J.beta, J.beta_zero, J.scale,
                  (int64_t)blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

template <class VT, class AT>
void launch_epilogue_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                           hipStream_t st) {
    hipLaunchKernelGGL((epilogue_batch_kernel<VT, AT>), dim3(nblocks), dim3(RED_THREADS), 0, st,
                       jobs, pre, nj);
    MSR_CHECK_LAUNCH();
}

template <class CT>
__global__ void scale_kernel(CT* __restrict__ x, int64_t n, CT beta, int beta_zero) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = beta_zero ? cz<CT>() : cmul(beta, x[i]);
}

template <class CT>
void launch_scale(CT* x, int64_t n, CT beta, bool beta_zero, hipStream_t st) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(scale_kernel<CT>, dim3(grid), dim3(256), 0, st, x, n, beta, (int)beta_zero);
    MSR_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ CG bodies
// r = b (- c); u = 0; x = 0 or x_in; partial ||r||^2
template <class VT, class AT>
__device__ __forceinline__ void cg_start_body(const float2* __restrict__ b,
                                              const float2* __restrict__ x_in,
                                              const CgVecs<VT, AT>& v, int64_t n, int zero_x,
                                              int r_minus_c, double* __restrict__ part, int blk,
                                              int nblk) {
    __shared__ double s[RED_THREADS / 64];
    double acc = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)nblk * blockDim.x) {
        const float2 bb = b[i];
        double rx = bb.x, ry = bb.y;
        if (r_minus_c) {
            const AT cc = v.c[i];
            rx -= (double)cc.x;
            ry -= (double)cc.y;
        }
        v.r[i] = CTraits<VT>::make(rx, ry);
        v.u[i] = cz<VT>();
        if (zero_x) v.x[i] = cz<VT>();
        else if ((const void*)x_in != (const void*)v.x) v.x[i] = cto<VT>(x_in[i]);
        // norm of the stored residual
        const VT rr = v.r[i];
        acc += (double)rr.x * rr.x + (double)rr.y * rr.y;
    }
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) part[blk] = t;
}

// Sum of nb block partials part[i * stride + off] by one wavefront in a fixed
// order (lane l takes i = l, l + 64, ..., then a shuffle tree); valid in lane 0.
// Called by all 64 lanes of the block.
__device__ __forceinline__ double wave_sum_partials(const double* __restrict__ part, int nb,
                                                    int stride, int off) {
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += 64) a += part[i * stride + off];
    return warp_sum(a);
}

__device__ __forceinline__ void cg_start_finish_body(const double* __restrict__ part, int nb,
                                                     CgScalars* S, double reltol, double abstol,
                                                     int maxiter) {
    const double a = wave_sum_partials(part, nb, 1, 0);
    if (threadIdx.x != 0) return;
    const double res = sqrt(a);
    S->res = res;
    S->prev = 1.0;
    S->tol = fmax(reltol * res, abstol);
    S->iter = 0;
    S->maxiter = maxiter;
    S->alpha_re = S->alpha_im = 0.0;
    S->done = (maxiter <= 0 || res <= S->tol) ? 1 : 0;
}

template <class VT, class AT>
__device__ __forceinline__ void cg_u_body(const CgVecs<VT, AT>& v, int64_t n,
                                          const CgScalars* __restrict__ S, int64_t blk,
                                          int64_t nblk) {
    if (S->done) return;
    using T = real_t<VT>;
    const T beta = (T)((S->res * S->res) / (S->prev * S->prev));
    for (int64_t i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
        const VT a = v.r[i], w = v.u[i];
        VT nu;
        nu.x = a.x + beta * w.x;
        nu.y = a.y + beta * w.y;
        v.u[i] = nu;
        if ((void*)v.ua != (void*)v.u) v.ua[i] = cto<AT>(nu);
    }
}

// c = scale sum_g partial_g; partial dot(u, c)
template <class VT, class AT>
__device__ __forceinline__ void cg_reduce_dot_body(const AT* __restrict__ partial, int G,
                                                   int64_t n, const CgVecs<VT, AT>& v,
                                                   double scale, double* __restrict__ part,
                                                   const CgScalars* __restrict__ S, int blk,
                                                   int nblk) {
    if (S->done) return;
    using RA = real_t<AT>;
    __shared__ double s[RED_THREADS / 64];
    double ar = 0.0, ai = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)nblk * blockDim.x) {
        AT c = partial[i];
        for (int g = 1; g < G; ++g) c = cadd(c, partial[(int64_t)g * n + i]);
        c = cscale(c, (RA)scale);
        v.c[i] = c;
        const VT w = v.u[i];
        // conj(u) * c
        ar += (double)w.x * c.x + (double)w.y * c.y;
        ai += (double)w.x * c.y - (double)w.y * c.x;
    }
    double tr = block_sum(ar, s);
    double ti = block_sum(ai, s);
    if (threadIdx.x == 0) {
        part[2 * blk] = tr;
        part[2 * blk + 1] = ti;
    }
}

__device__ __forceinline__ void cg_alpha_body(const double* __restrict__ part, int nb,
                                              CgScalars* S) {
    if (S->done) return;
    const double dr = wave_sum_partials(part, nb, 2, 0);
    const double di = wave_sum_partials(part, nb, 2, 1);
    if (threadIdx.x != 0) return;
    const double r2 = S->res * S->res;
    const double den = dr * dr + di * di;
    S->dot_re = dr;
    S->dot_im = di;
    S->alpha_re = r2 * dr / den;
    S->alpha_im = -r2 * di / den;
}

// x += alpha u; r -= alpha c; partial ||r||^2.  With dpart, alpha is formed
// here from the nb dot partials (as cg_alpha_body does: one wavefront sums them
// in a fixed order, so every block gets the same value) and block 0 stores it
// and the dot in Sw.
template <class VT, class AT>
__device__ __forceinline__ void cg_xr_body(const CgVecs<VT, AT>& v, int64_t n,
                                           double* __restrict__ part, const CgScalars* S,
                                           int blk, int nblk, const double* dpart = nullptr,
                                           int dnb = 0, CgScalars* Sw = nullptr) {
    if (S->done) return;
    using T = real_t<VT>;
    __shared__ double s[RED_THREADS / 64];
    double alr, ali;
    if (dpart) {
        __shared__ double s_alpha[2];
        if (threadIdx.x < 64) {
            const double dr = wave_sum_partials(dpart, dnb, 2, 0);
            const double di = wave_sum_partials(dpart, dnb, 2, 1);
            if (threadIdx.x == 0) {
                const double r2 = S->res * S->res;
                const double den = dr * dr + di * di;
                s_alpha[0] = r2 * dr / den;
                s_alpha[1] = -r2 * di / den;
                if (blk == 0) {
                    Sw->dot_re = dr;
                    Sw->dot_im = di;
                    Sw->alpha_re = s_alpha[0];
                    Sw->alpha_im = s_alpha[1];
                }
            }
        }
        __syncthreads();
        alr = s_alpha[0];
        ali = s_alpha[1];
    } else {
        alr = S->alpha_re;
        ali = S->alpha_im;
    }
    const T ar = (T)alr, ai = (T)ali;
    double acc = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)nblk * blockDim.x) {
        VT xv = v.x[i], rv = v.r[i];
        const VT u = v.u[i];
        const AT ca = v.c[i];
        const T cx = (T)ca.x, cy = (T)ca.y;
        xv.x += ar * u.x - ai * u.y;
        xv.y += ar * u.y + ai * u.x;
        rv.x -= ar * cx - ai * cy;
        rv.y -= ar * cy + ai * cx;
        v.x[i] = xv;
        v.r[i] = rv;
        acc += (double)rv.x * rv.x + (double)rv.y * rv.y;
    }
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) part[blk] = t;
}

__device__ __forceinline__ void cg_iter_finish_body(const double* __restrict__ part, int nb,
                                                    CgScalars* S) {
    if (S->done) return;
    const double a = wave_sum_partials(part, nb, 1, 0);
    if (threadIdx.x != 0) return;
    S->prev = S->res;
    S->res = sqrt(a);
    S->iter += 1;
    S->done = (S->iter >= S->maxiter || S->res <= S->tol) ? 1 : 0;
}

__device__ __forceinline__ void to_f32_body(const double2* __restrict__ a, float2* __restrict__ b,
                                            int64_t n, int64_t blk, int64_t nblk) {
    for (int64_t i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x)
        b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// ------------------------------------------------------------------ single-operator kernels
template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_start_kernel(const float2* __restrict__ b, const float2* __restrict__ x_in, CgVecs<VT, AT> v,
                int64_t n, int zero_x, int r_minus_c, double* __restrict__ part) {
    cg_start_body(b, x_in, v, n, zero_x, r_minus_c, part, blockIdx.x, gridDim.x);
}

__global__ void cg_start_finish(const double* __restrict__ part, int nb, CgScalars* S,
                                double reltol, double abstol, int maxiter) {
    cg_start_finish_body(part, nb, S, reltol, abstol, maxiter);
}

template <class VT, class AT>
void launch_cg_start(const float2* b, const float2* x_in, const CgVecs<VT, AT>& v, int64_t n,
                     bool zero_x, double* part, CgScalars* S, double reltol, double abstol,
                     int maxiter, bool r_minus_c, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL((cg_start_kernel<VT, AT>), dim3(g), dim3(RED_THREADS), 0, st, b, x_in, v,
                       n, (int)zero_x, (int)r_minus_c, part);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_start_finish, dim3(1), dim3(64), 0, st, part, g, S, reltol, abstol,
                       maxiter);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
__global__ void cg_u_kernel(CgVecs<VT, AT> v, int64_t n, const CgScalars* __restrict__ S) {
    cg_u_body(v, n, S, blockIdx.x, gridDim.x);
}

template <class VT, class AT>
void launch_cg_u(const CgVecs<VT, AT>& v, int64_t n, CgScalars* S, hipStream_t st) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL((cg_u_kernel<VT, AT>), dim3(grid), dim3(256), 0, st, v, n, S);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_reduce_dot_kernel(const AT* __restrict__ partial, int G, int64_t n, CgVecs<VT, AT> v,
                     double scale, double* __restrict__ part, const CgScalars* __restrict__ S) {
    cg_reduce_dot_body(partial, G, n, v, scale, part, S, blockIdx.x, gridDim.x);
}

__global__ void cg_alpha_kernel(const double* __restrict__ part, int nb, CgScalars* S) {
    cg_alpha_body(part, nb, S);
}

template <class VT, class AT>
void launch_cg_reduce_dot(const AT* partial, int G, int64_t n, const CgVecs<VT, AT>& v,
                          double scale, double* part, CgScalars* S, hipStream_t st,
                          bool step_kernel) {
    const int g = red_grid(n);
    hipLaunchKernelGGL((cg_reduce_dot_kernel<VT, AT>), dim3(g), dim3(RED_THREADS), 0, st, partial,
                       G, n, v, scale, part, S);
    MSR_CHECK_LAUNCH();
    if (!step_kernel) return;
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(64), 0, st, part, g, S);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_xr_kernel(CgVecs<VT, AT> v, int64_t n, double* __restrict__ part,
             const CgScalars* __restrict__ S) {
    cg_xr_body(v, n, part, S, blockIdx.x, gridDim.x);
}

__global__ void cg_iter_finish(const double* __restrict__ part, int nb, CgScalars* S) {
    cg_iter_finish_body(part, nb, S);
}

template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_xr_step_kernel(CgVecs<VT, AT> v, int64_t n, const double* __restrict__ dpart, int dnb,
                  double* __restrict__ part, CgScalars* S) {
    cg_xr_body(v, n, part, S, blockIdx.x, gridDim.x, dpart, dnb, S);
}

template <class VT, class AT>
void launch_cg_xr_step(const CgVecs<VT, AT>& v, int64_t n, double* part, CgScalars* S,
                       hipStream_t st) {
    const int g = red_grid(n);
    double* rpart = part + 2 * RED_BLOCKS;
    hipLaunchKernelGGL((cg_xr_step_kernel<VT, AT>), dim3(g), dim3(RED_THREADS), 0, st, v, n, part,
                       g, rpart, S);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_iter_finish, dim3(1), dim3(64), 0, st, rpart, g, S);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
void launch_cg_xr(const CgVecs<VT, AT>& v, int64_t n, double* part, CgScalars* S, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL((cg_xr_kernel<VT, AT>), dim3(g), dim3(RED_THREADS), 0, st, v, n, part, S);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_iter_finish, dim3(1), dim3(64), 0, st, part, g, S);
    MSR_CHECK_LAUNCH();
}

__global__ void to_f32_kernel(const double2* __restrict__ a, float2* __restrict__ b, int64_t n) {
    to_f32_body(a, b, n, blockIdx.x, gridDim.x);
}

void launch_to_f32(const double2* x64, float2* x32, int64_t n, hipStream_t st) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(to_f32_kernel, dim3(grid), dim3(256), 0, st, x64, x32, n);
    MSR_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ batched kernels
// Vector launches: nblocks = pre[nj] blocks of RED_THREADS threads; job j owns
// red_grid(n_j) of them (the single-operator reduction partition).  Scalar
// steps: one 64-thread block per job.
template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_start_batch_kernel(const CgJob<VT, AT>* __restrict__ jobs, const int* __restrict__ pre, int nj,
                      int zero_x, int r_minus_c) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT, AT>& J = jobs[j];
    cg_start_body(J.b, J.x_in, J.v, J.n, zero_x, r_minus_c, J.part, blockIdx.x - pre[j],
                  pre[j + 1] - pre[j]);
}

template <class VT, class AT>
__global__ void cg_start_finish_batch(const CgJob<VT, AT>* __restrict__ jobs,
                                      const int* __restrict__ pre, double reltol, double abstol,
                                      int maxiter) {
    const int j = blockIdx.x;
    cg_start_finish_body(jobs[j].part, pre[j + 1] - pre[j], jobs[j].S, reltol, abstol, maxiter);
}

template <class VT, class AT>
void launch_cg_start_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                           bool zero_x, bool r_minus_c, double reltol, double abstol, int maxiter,
                           hipStream_t st) {
    hipLaunchKernelGGL((cg_start_batch_kernel<VT, AT>), dim3(nblocks), dim3(RED_THREADS), 0, st,
                       jobs, pre, nj, (int)zero_x, (int)r_minus_c);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL((cg_start_finish_batch<VT, AT>), dim3(nj), dim3(64), 0, st, jobs, pre,
                       reltol, abstol, maxiter);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_u_batch_kernel(const CgJob<VT, AT>* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT, AT>& J = jobs[j];
    cg_u_body(J.v, J.n, J.S, (int64_t)blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

template <class VT, class AT>
void launch_cg_u_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                       hipStream_t st) {
    hipLaunchKernelGGL((cg_u_batch_kernel<VT, AT>), dim3(nblocks), dim3(RED_THREADS), 0, st, jobs,
                       pre, nj);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_reduce_dot_batch_kernel(const CgJob<VT, AT>* __restrict__ jobs, const int* __restrict__ pre,
                           int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT, AT>& J = jobs[j];
    cg_reduce_dot_body(J.partial, J.G, J.n, J.v, J.scale, J.part, J.S, blockIdx.x - pre[j],
                       pre[j + 1] - pre[j]);
}

template <class VT, class AT>
__global__ void cg_alpha_batch(const CgJob<VT, AT>* __restrict__ jobs, const int* __restrict__ pre) {
    const int j = blockIdx.x;
    cg_alpha_body(jobs[j].part, pre[j + 1] - pre[j], jobs[j].S);
}

template <class VT, class AT>
void launch_cg_reduce_dot_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                                hipStream_t st) {
    hipLaunchKernelGGL((cg_reduce_dot_batch_kernel<VT, AT>), dim3(nblocks), dim3(RED_THREADS), 0,
                       st, jobs, pre, nj);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL((cg_alpha_batch<VT, AT>), dim3(nj), dim3(64), 0, st, jobs, pre);
    MSR_CHECK_LAUNCH();
}

template <class VT, class AT>
__global__ void __launch_bounds__(RED_THREADS)
cg_xr_batch_kernel(const CgJob<VT, AT>* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT, AT>& J = jobs[j];
    cg_xr_body(J.v, J.n, J.part, J.S, blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

template <class VT, class AT>
__global__ void cg_iter_finish_batch(const CgJob<VT, AT>* __restrict__ jobs,
                                     const int* __restrict__ pre) {
    const int j = blockIdx.x;
    cg_iter_finish_body(jobs[j].part, pre[j + 1] - pre[j], jobs[j].S);
}

template <class VT, class AT>
void launch_cg_xr_batch(const CgJob<VT, AT>* jobs, const int* pre, int nj, int nblocks,
                        hipStream_t st) {
    hipLaunchKernelGGL((cg_xr_batch_kernel<VT, AT>), dim3(nblocks), dim3(RED_THREADS), 0, st, jobs,
                       pre, nj);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL((cg_iter_finish_batch<VT, AT>), dim3(nj), dim3(64), 0, st, jobs, pre);
    MSR_CHECK_LAUNCH();
}

template <class AT>
__global__ void __launch_bounds__(RED_THREADS)
to_f32_batch_kernel(const CgJob<double2, AT>* __restrict__ jobs, const int* __restrict__ pre,
                    int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<double2, AT>& J = jobs[j];
    to_f32_body(J.v.x, J.x_out, J.n, (int64_t)blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

template <class AT>
void launch_to_f32_batch(const CgJob<double2, AT>* jobs, const int* pre, int nj, int nblocks,
                         hipStream_t st) {
    hipLaunchKernelGGL(to_f32_batch_kernel<AT>, dim3(nblocks), dim3(RED_THREADS), 0, st, jobs, pre,
                       nj);
    MSR_CHECK_LAUNCH();
}

#define MSR_INST_CG(VT, AT)                                                                      \
    template void launch_cg_start<VT, AT>(const float2*, const float2*, const CgVecs<VT, AT>&,   \
                                          int64_t, bool, double*, CgScalars*, double, double,    \
                                          int, bool, hipStream_t);                               \
    template void launch_cg_u<VT, AT>(const CgVecs<VT, AT>&, int64_t, CgScalars*, hipStream_t);  \
    template void launch_cg_reduce_dot<VT, AT>(const AT*, int, int64_t, const CgVecs<VT, AT>&,   \
                                               double, double*, CgScalars*, hipStream_t, bool);  \
    template void launch_cg_xr<VT, AT>(const CgVecs<VT, AT>&, int64_t, double*, CgScalars*,      \
                                       hipStream_t);                                             \
    template void launch_cg_xr_step<VT, AT>(const CgVecs<VT, AT>&, int64_t, double*, CgScalars*, \
                                            hipStream_t);                                        \
    template void launch_cg_start_batch<VT, AT>(const CgJob<VT, AT>*, const int*, int, int,      \
                                                bool, bool, double, double, int, hipStream_t);   \
    template void launch_cg_u_batch<VT, AT>(const CgJob<VT, AT>*, const int*, int, int,          \
                                            hipStream_t);                                        \
    template void launch_cg_reduce_dot_batch<VT, AT>(const CgJob<VT, AT>*, const int*, int, int, \
                                                     hipStream_t);                               \
    template void launch_cg_xr_batch<VT, AT>(const CgJob<VT, AT>*, const int*, int, int,         \
                                             hipStream_t);                                       \
    template void launch_epilogue_batch<VT, AT>(const CgJob<VT, AT>*, const int*, int, int,      \
                                                hipStream_t);
MSR_INST_CG(float2, float2)
MSR_INST_CG(double2, float2)
MSR_INST_CG(double2, double2)
#undef MSR_INST_CG
template void launch_to_f32_batch<float2>(const CgJob<double2, float2>*, const int*, int, int,
                                          hipStream_t);
template void launch_to_f32_batch<double2>(const CgJob<double2, double2>*, const int*, int, int,
                                           hipStream_t);
template void launch_epilogue<float2>(const float2*, int, int64_t, float2*, float2, float2, bool,
                                      double, hipStream_t, const int*);
template void launch_epilogue<double2>(const double2*, int, int64_t, double2*, double2, double2,
                                       bool, double, hipStream_t, const int*);
template void launch_scale<float2>(float2*, int64_t, float2, bool, hipStream_t);
template void launch_scale<double2>(double2*, int64_t, double2, bool, hipStream_t);

}  // namespace msr
