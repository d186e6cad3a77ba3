// Vector kernels and the device CG recurrence with FP64 reductions.
//
// The CG follows IterativeSolvers.jl 0.9 (called by the reference at
// test/reconstruct_radial_gpu.jl:99 and src/CoilMaps.jl:136):
//   u = r + (res^2/prev^2) u;  c = A u;  alpha = res^2 / dot(u, c);
//   x += alpha u;  r -= alpha c;  prev = res;  res = ||r||
// stopping when iter >= maxiter or res <= max(reltol*||r0||, abstol).
// Dot products, norms and step lengths are accumulated and kept in FP64 on the
// device, reduced in a fixed order (block partials, then one block) so results
// are run-to-run deterministic.  The vectors are ComplexF32 or FP64 (VT).
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
__device__ __forceinline__ void epilogue_body(const float2* __restrict__ partial, int G, int64_t n,
                                              float2* __restrict__ x, float2 alpha, float2 beta,
                                              int beta_zero, float scale, int64_t blk,
                                              int64_t nblk) {
    for (int64_t i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
        float2 s = partial[i];
        for (int g = 1; g < G; ++g) s = cadd(s, partial[(int64_t)g * n + i]);
        s = cscale(s, scale);
        float2 v = cmul(alpha, s);
        if (!beta_zero) v = cadd(v, cmul(beta, x[i]));
        x[i] = v;
    }
}

__global__ void epilogue_kernel(const float2* __restrict__ partial, int G, int64_t n,
                                float2* __restrict__ x, float2 alpha, float2 beta, int beta_zero,
                                float scale, const int* skip) {
    if (skip && *skip) return;
    epilogue_body(partial, G, n, x, alpha, beta, beta_zero, scale, blockIdx.x, gridDim.x);
}

void launch_epilogue(const float2* partial, int G, int64_t n, float2* x, float2 alpha, float2 beta,
                     bool beta_zero, float scale, hipStream_t st, const int* skip) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(epilogue_kernel, dim3(grid), dim3(256), 0, st, partial, G, n, x, alpha,
                       beta, (int)beta_zero, scale, skip);
    MSR_CHECK_LAUNCH();
}

__global__ void epilogue_batch_kernel(const CgJob<float2>* __restrict__ jobs,
                                      const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<float2>& J = jobs[j];
    epilogue_body(J.partial, J.G, J.n, J.y, J.alpha, J.beta, J.beta_zero, J.scale,
                  (int64_t)blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

void launch_epilogue_batch(const CgJob<float2>* jobs, const int* pre, int nj, int nblocks,
                           hipStream_t st) {
    hipLaunchKernelGGL(epilogue_batch_kernel, dim3(nblocks), dim3(RED_THREADS), 0, st, jobs, pre, nj);
    MSR_CHECK_LAUNCH();
}

__global__ void scale_kernel(float2* __restrict__ x, int64_t n, float2 beta, int beta_zero) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = beta_zero ? make_float2(0.f, 0.f) : cmul(beta, x[i]);
}

void launch_scale(float2* x, int64_t n, float2 beta, bool beta_zero, hipStream_t st) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(scale_kernel, dim3(grid), dim3(256), 0, st, x, n, beta, (int)beta_zero);
    MSR_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ CG bodies
template <class VT> struct VecT;
template <> struct VecT<float2> {
    using S = float;
    __device__ static float2 make(double a, double b) { return make_float2((float)a, (float)b); }
};
template <> struct VecT<double2> {
    using S = double;
    __device__ static double2 make(double a, double b) { return make_double2(a, b); }
};

template <class VT>
__device__ __forceinline__ VT to_vt(float2 a) { return VecT<VT>::make(a.x, a.y); }

// r = b (- c); u = 0; x = 0 or x_in; partial ||r||^2
template <class VT>
__device__ __forceinline__ void cg_start_body(const float2* __restrict__ b,
                                              const float2* __restrict__ x_in, const CgVecs<VT>& v,
                                              int64_t n, int zero_x, int r_minus_c,
                                              double* __restrict__ part, int blk, int nblk) {
    __shared__ double s[RED_THREADS / 64];
    double acc = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)nblk * blockDim.x) {
        const float2 bb = b[i];
        double rx = bb.x, ry = bb.y;
        if (r_minus_c) {
            const float2 cc = v.c[i];
            rx -= cc.x;
            ry -= cc.y;
        }
        v.r[i] = VecT<VT>::make(rx, ry);
        v.u[i] = VecT<VT>::make(0.0, 0.0);
        if (zero_x) v.x[i] = VecT<VT>::make(0.0, 0.0);
        else if ((const void*)x_in != (const void*)v.x) v.x[i] = to_vt<VT>(x_in[i]);
        // norm of the stored residual
        const VT rr = v.r[i];
        acc += (double)rr.x * rr.x + (double)rr.y * rr.y;
    }
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) part[blk] = t;
}

__device__ __forceinline__ void cg_start_finish_body(const double* __restrict__ part, int nb,
                                                     CgScalars* S, double reltol, double abstol,
                                                     int maxiter) {
    if (threadIdx.x != 0) return;
    double a = 0.0;
    for (int i = 0; i < nb; ++i) a += part[i];
    const double res = sqrt(a);
    S->res = res;
    S->prev = 1.0;
    S->tol = fmax(reltol * res, abstol);
    S->iter = 0;
    S->maxiter = maxiter;
    S->alpha_re = S->alpha_im = 0.0;
    S->done = (maxiter <= 0 || res <= S->tol) ? 1 : 0;
}

template <class VT>
__device__ __forceinline__ void cg_u_body(const CgVecs<VT>& v, int64_t n,
                                          const CgScalars* __restrict__ S, int64_t blk,
                                          int64_t nblk) {
    if (S->done) return;
    using T = typename VecT<VT>::S;
    const T beta = (T)((S->res * S->res) / (S->prev * S->prev));
    for (int64_t i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
        const VT a = v.r[i], w = v.u[i];
        VT nu;
        nu.x = a.x + beta * w.x;
        nu.y = a.y + beta * w.y;
        v.u[i] = nu;
        if ((void*)v.u32 != (void*)v.u) v.u32[i] = make_float2((float)nu.x, (float)nu.y);
    }
}

// c = scale sum_g partial_g; partial dot(u, c)
template <class VT>
__device__ __forceinline__ void cg_reduce_dot_body(const float2* __restrict__ partial, int G,
                                                   int64_t n, const CgVecs<VT>& v, float scale,
                                                   double* __restrict__ part,
                                                   const CgScalars* __restrict__ S, int blk,
                                                   int nblk) {
    if (S->done) return;
    __shared__ double s[RED_THREADS / 64];
    double ar = 0.0, ai = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)nblk * blockDim.x) {
        float2 c = partial[i];
        for (int g = 1; g < G; ++g) c = cadd(c, partial[(int64_t)g * n + i]);
        c = cscale(c, scale);
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
    if (threadIdx.x != 0 || S->done) return;
    double dr = 0.0, di = 0.0;
    for (int i = 0; i < nb; ++i) {
        dr += part[2 * i];
        di += part[2 * i + 1];
    }
    const double r2 = S->res * S->res;
    const double den = dr * dr + di * di;
    S->dot_re = dr;
    S->dot_im = di;
    S->alpha_re = r2 * dr / den;
    S->alpha_im = -r2 * di / den;
}

// x += alpha u; r -= alpha c; partial ||r||^2
template <class VT>
__device__ __forceinline__ void cg_xr_body(const CgVecs<VT>& v, int64_t n,
                                           double* __restrict__ part,
                                           const CgScalars* __restrict__ S, int blk, int nblk) {
    if (S->done) return;
    using T = typename VecT<VT>::S;
    __shared__ double s[RED_THREADS / 64];
    const T ar = (T)S->alpha_re, ai = (T)S->alpha_im;
    double acc = 0.0;
    for (int64_t i = (int64_t)blk * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)nblk * blockDim.x) {
        VT xv = v.x[i], rv = v.r[i];
        const VT u = v.u[i];
        const float2 c32 = v.c[i];
        const T cx = (T)c32.x, cy = (T)c32.y;
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
    if (threadIdx.x != 0 || S->done) return;
    double a = 0.0;
    for (int i = 0; i < nb; ++i) a += part[i];
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
template <class VT>
__global__ void __launch_bounds__(RED_THREADS)
cg_start_kernel(const float2* __restrict__ b, const float2* __restrict__ x_in, CgVecs<VT> v,
                int64_t n, int zero_x, int r_minus_c, double* __restrict__ part) {
    cg_start_body(b, x_in, v, n, zero_x, r_minus_c, part, blockIdx.x, gridDim.x);
}

__global__ void cg_start_finish(const double* __restrict__ part, int nb, CgScalars* S,
                                double reltol, double abstol, int maxiter) {
    cg_start_finish_body(part, nb, S, reltol, abstol, maxiter);
}

template <class VT>
void launch_cg_start(const float2* b, const float2* x_in, const CgVecs<VT>& v, int64_t n,
                     bool zero_x, double* part, CgScalars* S, double reltol, double abstol,
                     int maxiter, bool r_minus_c, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(cg_start_kernel<VT>, dim3(g), dim3(RED_THREADS), 0, st, b, x_in, v, n,
                       (int)zero_x, (int)r_minus_c, part);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_start_finish, dim3(1), dim3(64), 0, st, part, g, S, reltol, abstol,
                       maxiter);
    MSR_CHECK_LAUNCH();
}

template <class VT>
__global__ void cg_u_kernel(CgVecs<VT> v, int64_t n, const CgScalars* __restrict__ S) {
    cg_u_body(v, n, S, blockIdx.x, gridDim.x);
}

template <class VT>
void launch_cg_u(const CgVecs<VT>& v, int64_t n, CgScalars* S, hipStream_t st) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(cg_u_kernel<VT>, dim3(grid), dim3(256), 0, st, v, n, S);
    MSR_CHECK_LAUNCH();
}

template <class VT>
__global__ void __launch_bounds__(RED_THREADS)
cg_reduce_dot_kernel(const float2* __restrict__ partial, int G, int64_t n, CgVecs<VT> v,
                     float scale, double* __restrict__ part, const CgScalars* __restrict__ S) {
    cg_reduce_dot_body(partial, G, n, v, scale, part, S, blockIdx.x, gridDim.x);
}

__global__ void cg_alpha_kernel(const double* __restrict__ part, int nb, CgScalars* S) {
    cg_alpha_body(part, nb, S);
}

template <class VT>
void launch_cg_reduce_dot(const float2* partial, int G, int64_t n, const CgVecs<VT>& v,
                          float scale, double* part, CgScalars* S, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(cg_reduce_dot_kernel<VT>, dim3(g), dim3(RED_THREADS), 0, st, partial, G,
                       n, v, scale, part, S);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(64), 0, st, part, g, S);
    MSR_CHECK_LAUNCH();
}

template <class VT>
__global__ void __launch_bounds__(RED_THREADS)
cg_xr_kernel(CgVecs<VT> v, int64_t n, double* __restrict__ part, const CgScalars* __restrict__ S) {
    cg_xr_body(v, n, part, S, blockIdx.x, gridDim.x);
}

__global__ void cg_iter_finish(const double* __restrict__ part, int nb, CgScalars* S) {
    cg_iter_finish_body(part, nb, S);
}

template <class VT>
void launch_cg_xr(const CgVecs<VT>& v, int64_t n, double* part, CgScalars* S, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(cg_xr_kernel<VT>, dim3(g), dim3(RED_THREADS), 0, st, v, n, part, S);
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
template <class VT>
__global__ void __launch_bounds__(RED_THREADS)
cg_start_batch_kernel(const CgJob<VT>* __restrict__ jobs, const int* __restrict__ pre, int nj,
                      int zero_x, int r_minus_c) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT>& J = jobs[j];
    cg_start_body(J.b, J.x_in, J.v, J.n, zero_x, r_minus_c, J.part, blockIdx.x - pre[j],
                  pre[j + 1] - pre[j]);
}

template <class VT>
__global__ void cg_start_finish_batch(const CgJob<VT>* __restrict__ jobs,
                                      const int* __restrict__ pre, double reltol, double abstol,
                                      int maxiter) {
    const int j = blockIdx.x;
    cg_start_finish_body(jobs[j].part, pre[j + 1] - pre[j], jobs[j].S, reltol, abstol, maxiter);
}

template <class VT>
void launch_cg_start_batch(const CgJob<VT>* jobs, const int* pre, int nj, int nblocks,
                           bool zero_x, bool r_minus_c, double reltol, double abstol, int maxiter,
                           hipStream_t st) {
    hipLaunchKernelGGL(cg_start_batch_kernel<VT>, dim3(nblocks), dim3(RED_THREADS), 0, st, jobs,
                       pre, nj, (int)zero_x, (int)r_minus_c);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_start_finish_batch<VT>, dim3(nj), dim3(64), 0, st, jobs, pre, reltol,
                       abstol, maxiter);
    MSR_CHECK_LAUNCH();
}

template <class VT>
__global__ void __launch_bounds__(RED_THREADS)
cg_u_batch_kernel(const CgJob<VT>* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT>& J = jobs[j];
    cg_u_body(J.v, J.n, J.S, (int64_t)blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

template <class VT>
void launch_cg_u_batch(const CgJob<VT>* jobs, const int* pre, int nj, int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(cg_u_batch_kernel<VT>, dim3(nblocks), dim3(RED_THREADS), 0, st, jobs, pre, nj);
    MSR_CHECK_LAUNCH();
}

template <class VT>
__global__ void __launch_bounds__(RED_THREADS)
cg_reduce_dot_batch_kernel(const CgJob<VT>* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT>& J = jobs[j];
    cg_reduce_dot_body(J.partial, J.G, J.n, J.v, J.scale, J.part, J.S, blockIdx.x - pre[j],
                       pre[j + 1] - pre[j]);
}

template <class VT>
__global__ void cg_alpha_batch(const CgJob<VT>* __restrict__ jobs, const int* __restrict__ pre) {
    const int j = blockIdx.x;
    cg_alpha_body(jobs[j].part, pre[j + 1] - pre[j], jobs[j].S);
}

template <class VT>
void launch_cg_reduce_dot_batch(const CgJob<VT>* jobs, const int* pre, int nj, int nblocks,
                                hipStream_t st) {
    hipLaunchKernelGGL(cg_reduce_dot_batch_kernel<VT>, dim3(nblocks), dim3(RED_THREADS), 0, st,
                       jobs, pre, nj);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_alpha_batch<VT>, dim3(nj), dim3(64), 0, st, jobs, pre);
    MSR_CHECK_LAUNCH();
}

template <class VT>
__global__ void __launch_bounds__(RED_THREADS)
cg_xr_batch_kernel(const CgJob<VT>* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<VT>& J = jobs[j];
    cg_xr_body(J.v, J.n, J.part, J.S, blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

template <class VT>
__global__ void cg_iter_finish_batch(const CgJob<VT>* __restrict__ jobs, const int* __restrict__ pre) {
    const int j = blockIdx.x;
    cg_iter_finish_body(jobs[j].part, pre[j + 1] - pre[j], jobs[j].S);
}

template <class VT>
void launch_cg_xr_batch(const CgJob<VT>* jobs, const int* pre, int nj, int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(cg_xr_batch_kernel<VT>, dim3(nblocks), dim3(RED_THREADS), 0, st, jobs, pre, nj);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_iter_finish_batch<VT>, dim3(nj), dim3(64), 0, st, jobs, pre);
    MSR_CHECK_LAUNCH();
}

__global__ void __launch_bounds__(RED_THREADS)
to_f32_batch_kernel(const CgJob<double2>* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    const CgJob<double2>& J = jobs[j];
    to_f32_body(J.v.x, J.x_out, J.n, (int64_t)blockIdx.x - pre[j], pre[j + 1] - pre[j]);
}

void launch_to_f32_batch(const CgJob<double2>* jobs, const int* pre, int nj, int nblocks,
                         hipStream_t st) {
    hipLaunchKernelGGL(to_f32_batch_kernel, dim3(nblocks), dim3(RED_THREADS), 0, st, jobs, pre, nj);
    MSR_CHECK_LAUNCH();
}

#define MSR_INST_CG(VT)                                                                          \
    template void launch_cg_start<VT>(const float2*, const float2*, const CgVecs<VT>&, int64_t,  \
                                      bool, double*, CgScalars*, double, double, int, bool,      \
                                      hipStream_t);                                              \
    template void launch_cg_u<VT>(const CgVecs<VT>&, int64_t, CgScalars*, hipStream_t);          \
    template void launch_cg_reduce_dot<VT>(const float2*, int, int64_t, const CgVecs<VT>&, float, \
                                           double*, CgScalars*, hipStream_t);                    \
    template void launch_cg_xr<VT>(const CgVecs<VT>&, int64_t, double*, CgScalars*, hipStream_t); \
    template void launch_cg_start_batch<VT>(const CgJob<VT>*, const int*, int, int, bool, bool,  \
                                            double, double, int, hipStream_t);                   \
    template void launch_cg_u_batch<VT>(const CgJob<VT>*, const int*, int, int, hipStream_t);    \
    template void launch_cg_reduce_dot_batch<VT>(const CgJob<VT>*, const int*, int, int,         \
                                                 hipStream_t);                                   \
    template void launch_cg_xr_batch<VT>(const CgJob<VT>*, const int*, int, int, hipStream_t);
MSR_INST_CG(float2)
MSR_INST_CG(double2)
#undef MSR_INST_CG

}  // namespace msr
