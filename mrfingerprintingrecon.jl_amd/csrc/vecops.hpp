// Vector kernels: apply epilogue (group reduction with alpha/beta), and the
// device conjugate-gradient steps with FP64 reductions.
#pragma once

#include "internal.hpp"

namespace msr {

constexpr int RED_BLOCKS = 512;   // blocks of the two-level FP64 reductions
constexpr int RED_THREADS = 256;

// x = beta*x + alpha*scale*sum_g partial[g]   (beta_zero: overwrite)
void launch_epilogue(const float2* partial, int G, int64_t n, float2* x, float2 alpha, float2 beta,
                     bool beta_zero, float scale, hipStream_t st, const int* skip);

// x = beta*x (or 0 when beta_zero)
void launch_scale(float2* x, int64_t n, float2 beta, bool beta_zero, hipStream_t st);

// CG building blocks (scalars in CgScalars on the device)
void launch_cg_start(const float2* b, float2* x, float2* r, float2* u, int64_t n, bool zero_x,
                     double* part, CgScalars* S, double reltol, double abstol, int maxiter,
                     bool r_minus_c, const float2* c, hipStream_t st);
void launch_cg_u(const float2* r, float2* u, int64_t n, CgScalars* S, hipStream_t st);
// c = scale*sum_g partial[g]; dot(u, c) partials; alpha = res^2 / dot
void launch_cg_reduce_dot(const float2* partial, int G, int64_t n, float2* c, const float2* u,
                          float scale, double* part, CgScalars* S, hipStream_t st);
void launch_cg_xr(float2* x, float2* r, const float2* u, const float2* c, int64_t n, double* part,
                  CgScalars* S, hipStream_t st);

}  // namespace msr
