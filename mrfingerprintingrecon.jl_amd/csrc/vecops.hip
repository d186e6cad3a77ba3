// Vector kernels and the device CG recurrence with FP64 reductions.
//
// The CG follows IterativeSolvers.jl 0.9 (called by the reference at
// test/reconstruct_radial_gpu.jl:99 and src/CoilMaps.jl:136):
//   u = r + (res^2/prev^2) u;  c = A u;  alpha = res^2 / dot(u, c);
//   x += alpha u;  r -= alpha c;  prev = res;  res = ||r||
// stopping when iter >= maxiter or res <= max(reltol*||r0||, abstol).
// Vectors stay FP32; dot products, norms and step lengths are accumulated and
// kept in FP64 on the device, reduced in a fixed order (block partials, then
// one block) so results are run-to-run deterministic.
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

static inline int red_grid(int64_t n) {
    int64_t g = (n + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
    if (g < 1) g = 1;
    if (g > RED_BLOCKS) g = RED_BLOCKS;
    return (int)g;
}

// ------------------------------------------------------------------ epilogue
__global__ void epilogue_kernel(const float2* __restrict__ partial, int G, int64_t n,
                                float2* __restrict__ x, float2 alpha, float2 beta, int beta_zero,
                                float scale, const int* skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float2 s = partial[i];
        for (int g = 1; g < G; ++g) s = cadd(s, partial[(int64_t)g * n + i]);
        s = cscale(s, scale);
        float2 v = cmul(alpha, s);
        if (!beta_zero) v = cadd(v, cmul(beta, x[i]));
        x[i] = v;
    }
}

void launch_epilogue(const float2* partial, int G, int64_t n, float2* x, float2 alpha, float2 beta,
                     bool beta_zero, float scale, hipStream_t st, const int* skip) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(epilogue_kernel, dim3(grid), dim3(256), 0, st, partial, G, n, x, alpha,
                       beta, (int)beta_zero, scale, skip);
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

// ------------------------------------------------------------------ CG
__global__ void __launch_bounds__(RED_THREADS)
cg_start_kernel(const float2* __restrict__ b, float2* __restrict__ x, float2* __restrict__ r,
                float2* __restrict__ u, int64_t n, int zero_x, int r_minus_c,
                const float2* __restrict__ c, double* __restrict__ part) {
    __shared__ double s[RED_THREADS / 64];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float2 v = b[i];
        if (r_minus_c) v = csub(v, c[i]);
        r[i] = v;
        u[i] = make_float2(0.f, 0.f);
        if (zero_x) x[i] = make_float2(0.f, 0.f);
        acc += (double)v.x * v.x + (double)v.y * v.y;
    }
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void cg_start_finish(const double* __restrict__ part, int nb, CgScalars* S,
                                double reltol, double abstol, int maxiter) {
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

void launch_cg_start(const float2* b, float2* x, float2* r, float2* u, int64_t n, bool zero_x,
                     double* part, CgScalars* S, double reltol, double abstol, int maxiter,
                     bool r_minus_c, const float2* c, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(cg_start_kernel, dim3(g), dim3(RED_THREADS), 0, st, b, x, r, u, n,
                       (int)zero_x, (int)r_minus_c, c, part);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_start_finish, dim3(1), dim3(64), 0, st, part, g, S, reltol, abstol,
                       maxiter);
    MSR_CHECK_LAUNCH();
}

__global__ void cg_u_kernel(const float2* __restrict__ r, float2* __restrict__ u, int64_t n,
                            const CgScalars* __restrict__ S) {
    if (S->done) return;
    const float beta = (float)((S->res * S->res) / (S->prev * S->prev));
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float2 a = r[i], v = u[i];
        u[i] = make_float2(fmaf(beta, v.x, a.x), fmaf(beta, v.y, a.y));
    }
}

void launch_cg_u(const float2* r, float2* u, int64_t n, CgScalars* S, hipStream_t st) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(cg_u_kernel, dim3(grid), dim3(256), 0, st, r, u, n, S);
    MSR_CHECK_LAUNCH();
}

__global__ void __launch_bounds__(RED_THREADS)
cg_reduce_dot_kernel(const float2* __restrict__ partial, int G, int64_t n, float2* __restrict__ c,
                     const float2* __restrict__ u, float scale, double* __restrict__ part,
                     const CgScalars* __restrict__ S) {
    if (S->done) return;
    __shared__ double s[RED_THREADS / 64];
    double ar = 0.0, ai = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float2 v = partial[i];
        for (int g = 1; g < G; ++g) v = cadd(v, partial[(int64_t)g * n + i]);
        v = cscale(v, scale);
        c[i] = v;
        const float2 w = u[i];
        // conj(u) * c
        ar += (double)w.x * v.x + (double)w.y * v.y;
        ai += (double)w.x * v.y - (double)w.y * v.x;
    }
    double tr = block_sum(ar, s);
    double ti = block_sum(ai, s);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = tr;
        part[2 * blockIdx.x + 1] = ti;
    }
}

__global__ void cg_alpha_kernel(const double* __restrict__ part, int nb, CgScalars* S) {
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

void launch_cg_reduce_dot(const float2* partial, int G, int64_t n, float2* c, const float2* u,
                          float scale, double* part, CgScalars* S, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(cg_reduce_dot_kernel, dim3(g), dim3(RED_THREADS), 0, st, partial, G, n, c,
                       u, scale, part, S);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(64), 0, st, part, g, S);
    MSR_CHECK_LAUNCH();
}

__global__ void __launch_bounds__(RED_THREADS)
cg_xr_kernel(float2* __restrict__ x, float2* __restrict__ r, const float2* __restrict__ u,
             const float2* __restrict__ c, int64_t n, double* __restrict__ part,
             const CgScalars* __restrict__ S) {
    if (S->done) return;
    __shared__ double s[RED_THREADS / 64];
    const float2 al = make_float2((float)S->alpha_re, (float)S->alpha_im);
    double acc = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        float2 xv = x[i], rv = r[i];
        cfma(xv, al, u[i]);
        const float2 ac = cmul(al, c[i]);
        rv = csub(rv, ac);
        x[i] = xv;
        r[i] = rv;
        acc += (double)rv.x * rv.x + (double)rv.y * rv.y;
    }
    double t = block_sum(acc, s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void cg_iter_finish(const double* __restrict__ part, int nb, CgScalars* S) {
    if (threadIdx.x != 0 || S->done) return;
    double a = 0.0;
    for (int i = 0; i < nb; ++i) a += part[i];
    S->prev = S->res;
    S->res = sqrt(a);
    S->iter += 1;
    S->done = (S->iter >= S->maxiter || S->res <= S->tol) ? 1 : 0;
}

void launch_cg_xr(float2* x, float2* r, const float2* u, const float2* c, int64_t n, double* part,
                  CgScalars* S, hipStream_t st) {
    const int g = red_grid(n);
    hipLaunchKernelGGL(cg_xr_kernel, dim3(g), dim3(RED_THREADS), 0, st, x, r, u, c, n, part, S);
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(cg_iter_finish, dim3(1), dim3(64), 0, st, part, g, S);
    MSR_CHECK_LAUNCH();
}

}  // namespace msr
