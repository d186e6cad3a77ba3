// Normal-operator internals shared by the API layer and the batch engine.
#pragma once

#include "apply2d.hpp"
#include <vector>

#include "internal.hpp"
#include "vecops.hpp"

namespace msr {

msr_normalop* op_new(msr_ctx* ctx, int d, const int64_t* N, int K, int lam_real, int cart = 0);
// Cartesian operator passes (cartesian.hip): partial <- unscaled sum_c conj(S_c)
// IFFT(Λ . FFT(S_c b)) on the N grid
void launch_cart_passes(msr_normalop* op, const void* b, bool f64, hipStream_t st, const int* skip);
// General Toeplitz passes on the line kernels (apply_lines.hip), for the
// shapes and coefficient counts outside the fused passes: partial <- unscaled
// sum over coils c_begin..c_end-1 (c_end < 0: all) of conj(S_c) crop(IFFT(
// Lambda . FFT(pad(S_c b)))), G = 1.
void launch_lines_passes(msr_normalop* op, const void* b, bool f64, hipStream_t st,
                         const int* skip, int c_begin = 0, int c_end = -1);
int lines_chunk_coils(const msr_normalop* op);
size_t lines_scratch_bytes(const msr_normalop* op, size_t es);
// the largest coefficient count (the fused passes take up to 8)
constexpr int MAX_COEFFS = 64;
void op_set_cmaps(msr_normalop* op, int ncoils, const void* const* cmaps, hipStream_t st);
void op_load_packed(msr_normalop* op, const void* lambda, const int64_t* kmask, int64_t nkm,
                    hipStream_t st);
void op_alloc_scratch(msr_normalop* op);
// tuned fused-2D knobs of the operator's shape on its context (tune.hip)
bool tuned_knobs(const msr_normalop* op, int am, ApplyKnobs* out);
void remember_knobs(const msr_normalop* op, int am, const ApplyKnobs& k);
// Measures the fused 2D apply of mode am under every knob set (Gb in 1, 2, 4,
// 8 up to C; both block orders of passes A and C), nrep applies each, checks
// that every set gives the first one's partial sums bit for bit, and keeps
// the fastest (also for later operators of the same shape on the context).
// times[TUNE_CONFIGS] (ms per apply, NaN where Gb > C; may be null), *chosen
// = 4 g + 2 order_a + order_c with Gb = 1 << g, or -1 for an operator without
// fused 2D passes.  nrep = 0 only reports the current knobs.
constexpr int TUNE_CONFIGS = 16;
void op_tune(msr_normalop* op, int am, int nrep, double* times, int* chosen, hipStream_t st);
// the context's tuned knobs to / from a text file (tune.hip); load returns
// the number of entries and refuses a file measured on another device type
void tune_save(msr_ctx* ctx, const char* path);
int tune_load(msr_ctx* ctx, const char* path);
void op_apply(msr_normalop* op, float2* x, const float2* b, float2 alpha, float2 beta,
              bool beta_zero, hipStream_t st);
// ComplexF64 apply; am = AM_F64 or AM_F64C (compact intermediates, 2D only)
void op_apply_f64(msr_normalop* op, double2* x, const double2* b, double2 alpha, double2 beta,
                  bool beta_zero, hipStream_t st, int am = AM_F64);
void op_apply_percoil(msr_normalop* op, float2* x, const float2* b, float2 alpha, float2 beta,
                      bool beta_zero, hipStream_t st);
void op_export(const msr_normalop* op, void* lambda_out, int64_t* kmask_out, hipStream_t st);
void op_cg(msr_normalop* op, float2* x, const float2* b, int maxiter, double reltol,
           double abstol, bool initially_zero, int mode, int prec, int* iters_out,
           double* res_out, hipStream_t st);

// Device-mode CG whose vectors, scalars and partials live in one static device
// array per device (msr_cg_noalloc): no allocation of any kind.
constexpr size_t CG_NOALLOC_BYTES = MSR_CG_NOALLOC_BYTES;  // include/msr.h
void cg_noalloc_setup(int device);  // once per device, with the device current
size_t cg_noalloc_bytes(int64_t n, int prec);
void op_cg_noalloc(msr_normalop* op, float2* x, const float2* b, int maxiter, double reltol,
                   double abstol, bool initially_zero, int prec, int* iters_out, double* res_out,
                   hipStream_t st);

// pieces used by the batch engine
// apply mode (ApplyMode) of a CG precision (MSR_PREC_*)
inline int apply_mode(int prec) {
    return prec == MSR_PREC_F64 ? AM_F64 : prec == MSR_PREC_F64C ? AM_F64C : AM_F32;
}
inline bool prec_valid(int prec) {
    return prec == MSR_PREC_F32 || prec == MSR_PREC_MIXED || prec == MSR_PREC_F64 ||
           prec == MSR_PREC_F64C;
}
Apply2DArgs args2d(msr_normalop* op, int am);          // b = skip = nullptr
void op_ensure_scratch(msr_normalop* op, bool f64);    // grow-only; not during capture
void op_passes(msr_normalop* op, const void* b, int am, hipStream_t st, const int* skip);
double op_scale(const msr_normalop* op);               // 1 / Toeplitz grid cells
void op_cg_buffers(msr_normalop* op, int prec);        // allocates; not during capture
template <class VT, class AT>
CgVecs<VT, AT> op_cg_vecs(msr_normalop* op, float2* x);  // x: caller's FP32 solution
std::vector<const void*> op_graph_key(const msr_normalop* op);  // buffer addresses
// Device bytes of the operator's own state for a solve at precision prec (coil
// maps, apply scratch and CG vectors; the kernel, shared with
// msr_normalop_create_shared, is not counted).
size_t op_footprint(const msr_normalop* op, int prec);
// Device bytes of the kernel and kmask.
size_t op_kernel_bytes(const msr_normalop* op);
// A new operator on an existing kernel and kmask (kcache.hip).
msr_normalop* op_on_kernel(msr_ctx* ctx, int d, const int64_t* N, int K, int lam_real,
                           int64_t nkm, const std::shared_ptr<DevBuf>& lam,
                           const std::shared_ptr<DevBuf>& kmask, int ncoils,
                           const void* const* cmaps, hipStream_t st, int cart = 0);

}  // namespace msr
