/*
 * msr.h -- C ABI of the MI355X (gfx950) Toeplitz subspace normal operator,
 * subspace backprojection and device conjugate gradient.
 *
 * Drop-in boundary for MRISubspaceRecon.jl (repo MagneticResonanceImaging/
 * MRFingerprintingRecon.jl): every entry point below names the reference
 * function it replaces.  The reference binds its GPU path by multiple dispatch
 * on CuArray inside a package extension (ext/MRISubspaceReconCUDAExt); an AMD
 * extension binds these symbols with `ccall` instead.
 *
 * Conventions
 *  - Arrays are device pointers in Julia column-major layout with the
 *    reference's shapes: trj (d, Nsamp, Nt) Float32, sample_mask (Nsamp, Nt)
 *    Bool (1 byte), data (Nsamp, Nt, C) ComplexF32, U (Nt, K) ComplexF32 or
 *    Float32, images (N1, N2[, N3], K) ComplexF32, cmaps one (N1, N2[, N3])
 *    ComplexF32 array per coil, packed kernel (P, Nkm) with P = K(K+1)/2
 *    (ComplexF32 for complex U, Float32 for real U), kmask Int64 1-based.
 *  - Complex numbers are interleaved float pairs (ComplexF32 == float2).
 *  - Every call is ordered on the given HIP stream (hipStream_t passed as
 *    void*, NULL = default stream).  No call synchronises the host unless its
 *    comment says so.
 *  - Return value: MSR_OK or an error code; msr_last_error() returns the
 *    message of the calling thread's last error.  Error codes map to the Julia
 *    exceptions the reference throws (ArgumentError, AssertionError,
 *    ErrorException).
 *  - Handles are not re-entrant: one in-flight call per operator handle, as
 *    the reference's operator shares its scratch arrays
 *    (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:36-37).
 *  - A trajectory handle may be shared: calls that use it from several
 *    threads are serialised on the host, and a backprojection on another
 *    stream than the handle's previous backprojection is ordered after it
 *    (the handle's sorted-sample plans and backprojection workspace are
 *    shared by all its calls).
 */
#ifndef MSR_H
#define MSR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSR_OK 0
#define MSR_ERR_ARGUMENT 1    /* ArgumentError (src/BackProjection.jl:183-202) */
#define MSR_ERR_ASSERTION 2   /* AssertionError (src/NFFTNormalOp.jl:59-61,107,128) */
#define MSR_ERR_ERROR 3       /* ErrorException (src/BackProjection.jl:179) */
#define MSR_ERR_UNSUPPORTED 4 /* size or option this build has no kernel for */
#define MSR_ERR_NOMEM 5       /* device allocation failed */
#define MSR_ERR_HIP 6         /* HIP runtime error */
#define MSR_ERR_IO 7          /* kernel file missing, unreadable, truncated or corrupt */

/* density_compensation modes (src/BackProjection.jl:169-181) */
#define MSR_DCF_NONE 0
#define MSR_DCF_RADIAL_2D 1
#define MSR_DCF_RADIAL_3D 2
#define MSR_DCF_ARRAY 3

/* conjugate-gradient execution modes */
#define MSR_CG_DEVICE 0 /* scalars stay on the device (FP64), no host sync per iteration */
#define MSR_CG_HOST 1   /* IterativeSolvers-style: dot/norm read back to the host each iteration */
#define MSR_CG_GRAPH 2  /* as MSR_CG_DEVICE, iterations captured once into a hipGraph and replayed */

/* CG precision (the `precision` argument of msr_cg / msr_batch_cg*) */
#define MSR_PREC_F32 0   /* ComplexF32 vectors and apply (the reference's arithmetic) */
#define MSR_PREC_MIXED 1 /* FP64 vectors, ComplexF32 apply on an FP32 copy of the search vector */
#define MSR_PREC_F64 2   /* FP64 vectors and ComplexF64 apply (FFTs, kernel multiply and
                            intermediates in FP64; kernel and coil maps stored as ComplexF32) */
#define MSR_PREC_F64C 3  /* as MSR_PREC_F64, with the fused 2D and 3D applies' FFT
                            intermediates stored in 8 bytes per complex value (a shared
                            exponent and two 28-bit mantissas: relative rounding 2^-27, 8x
                            below ComplexF32 storage); Cartesian operators and the
                            line-kernel apply keep ComplexF64 intermediates */

typedef struct msr_ctx msr_ctx;
typedef struct msr_traj msr_traj;
typedef struct msr_normalop msr_normalop;
typedef struct msr_batch msr_batch;

/* ------------------------------------------------------------------ misc */
/* Copies the calling thread's last error message into buf (NUL-terminated,
 * truncated to len); returns the full message length. */
size_t msr_last_error(char *buf, size_t len);
int msr_version(void);

/* Library context on one HIP device (twiddle tables, launch parameters).
 * No reference counterpart: CUDA.jl's implicit context. */
int msr_ctx_create(int device, msr_ctx **out);
int msr_ctx_destroy(msr_ctx *ctx);

/* --------------------------------------------------------- trajectories */
/* Sample selection shared by the kernel build and the backprojection:
 * trj[:, sample_mask] and the per-frame sample counts / offsets
 * (src/NFFTNormalOp.jl:126-131; ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:95-97;
 * ext/MRISubspaceReconCUDAExt/BackProjection.jl:37-48).  trj: (d, nsamp, nt)
 * Float32 device; mask: (nsamp, nt) uint8 device or NULL (all samples).
 * Synchronises the host once (the masked sample count is a host value, as in
 * the reference's sum(nsamp_t)).  AssertionError if the mask removes all
 * samples (src/NFFTNormalOp.jl:128). */
int msr_traj_create(msr_ctx *ctx, int d, int64_t nsamp, int64_t nt, const float *trj,
                    const uint8_t *mask, void *stream, msr_traj **out);
int msr_traj_info(const msr_traj *traj, int64_t *n_masked, int64_t *nt);
int msr_traj_destroy(msr_traj *traj);

/* ------------------------------------------------------ normal operator */
/* NFFTNormalOp(img_shape, trj, U; cmaps, sample_mask)
 * (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:1-13, which calls
 * calculate_kernel_noncartesian :89-191 and calculate_kmask_indcs :75-86).
 * U: (Nt, K) device, ComplexF32 (u_real = 0) or Float32 (u_real = 1; builds
 * the real symmetric kernel).  cmaps: host array of ncoils device pointers, or
 * NULL with ncoils = 0 for the reference default cmaps = (1,).
 * AssertionError if a grid size 2N is not 2,3,5-smooth
 * (src/NFFTNormalOp.jl:107).  Supported: every 2,3,5-smooth 2N up to 2048
 * (the kernel construction's 4N up to 2048, so N up to 512 there) and K up to
 * 64; MSR_ERR_UNSUPPORTED beyond.  The applies run fused passes for 2N in
 * {16, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512} along
 * the axes they transform in LDS (x and y in 2D, x and z in 3D) and K <= 8,
 * and the line-kernel passes otherwise (same results to rounding). */
int msr_normalop_create_from_traj(msr_ctx *ctx, const msr_traj *traj, int d,
                                  const int64_t *img_shape, int K, const void *U, int u_real,
                                  int ncoils, const void *const *cmaps, void *stream,
                                  msr_normalop **out);

/* NFFTNormalOp(img_shape, Λ, kmask_indcs; cmaps)
 * (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:22-70): lambda is the packed
 * (P, Nkm) kernel on the device, kmask the 1-based Int64 indices into the 2N
 * grid (device).  K is derived from P as (isqrt(8P+1)-1)/2 (:33).
 * AssertionError if a kmask entry is outside 1..prod(2N) (:28-30). */
int msr_normalop_create(msr_ctx *ctx, int d, const int64_t *img_shape, int64_t P,
                        int lambda_real, const void *lambda, const int64_t *kmask, int64_t nkm,
                        int ncoils, const void *const *cmaps, void *stream, msr_normalop **out);

/* Sizes of the operator's packed kernel (op.prod!.A.Λ and .kmask_indcs of the
 * reference, test/reconstruct_radial_gpu.jl:104). */
int msr_normalop_info(const msr_normalop *op, int *K, int64_t *P, int64_t *nkm, int *lambda_real);

/* Writes the packed kernel (P, Nkm) and the 1-based kmask into caller device
 * buffers of the sizes msr_normalop_info reports. */
int msr_normalop_export_kernel(const msr_normalop *op, void *lambda_out, int64_t *kmask_out,
                               void *stream);

/* mul!(x, ::_NFFTNormalOp, b, α, β) (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:193-218):
 * x = β x + α AᴴA b with x, b (N..., K) ComplexF32 on the device; β == 0
 * overwrites x without reading it (NaN in x is ignored, :195-199). */
int msr_normalop_apply(msr_normalop *op, void *x, const void *b, double alpha_re,
                       double alpha_im, double beta_re, double beta_im, void *stream);

/* The reference's GPU apply order, one coil at a time (fill, pad, forward FFT,
 * kernel multiply re-reading Λ, inverse FFT, crop-accumulate per coil, :204-216),
 * on this library's kernels.  Same result as msr_normalop_apply; kept as the
 * plain one-recon-at-a-time GPU path the fused apply is measured against. */
int msr_normalop_apply_percoil(msr_normalop *op, void *x, const void *b, double alpha_re,
                               double alpha_im, double beta_re, double beta_im, void *stream);

int msr_normalop_destroy(msr_normalop *op);

/* --------------------------------------------------------- backprojection */
/* calculate_backprojection(data, trj, cmaps; U, sample_mask, density_compensation)
 * (ext/MRISubspaceReconCUDAExt/BackProjection.jl:36-71; per-coil variant :1-34
 * when cmaps == NULL, which takes img_shape and writes (N..., K, C)).
 * data: (nsamp, nt, ncoils) ComplexF32 device (the full, unmasked array);
 * U: (Nt, K) device or NULL for the reference default I(Nt) (then K = Nt).
 * dcf_mode: MSR_DCF_*; dcf: masked-sample weights (Float32, length n_masked)
 * for MSR_DCF_ARRAY.  out: (N..., K) or (N..., K, C) ComplexF32 device.
 * ArgumentError on shape mismatch (src/BackProjection.jl:183-202);
 * ErrorException on an unknown dcf_mode (:179). */
int msr_backprojection(msr_ctx *ctx, const msr_traj *traj, int d, const int64_t *img_shape,
                       const void *data, int ncoils, int K, const void *U, int u_real,
                       const void *const *cmaps, int dcf_mode, const float *dcf, void *out,
                       void *stream);

/* msr_backprojection with coil maps, computed in the reference's call
 * sequence (ext/MRISubspaceReconCUDAExt/BackProjection.jl:56-66): one
 * single-channel type-1 NUFFT per (coefficient, coil), each spreading every
 * sample, accumulated with conj(S_c).  Same result; it exists to time the
 * reference's sequence on this library's kernels.  U and cmaps are required. */
int msr_backprojection_channelwise(msr_ctx *ctx, const msr_traj *traj, int d,
                                   const int64_t *img_shape, const void *data, int ncoils, int K,
                                   const void *U, int u_real, const void *const *cmaps,
                                   int dcf_mode, const float *dcf, void *out, void *stream);

/* --------------------------------------------------------- forward model */
/* The tests' data simulation (test/reconstruct_radial.jl:62-74: per coil and
 * frame, NonuniformFFTs exec_type2! of S_c * sum_k U[t,k] x_k), i.e. the
 * adjoint of msr_backprojection without density compensation:
 *   y_{j,c} = sum_k U[t_j,k] sum_n S_c[n] x_k[n] exp(-2 pi i k_j.(n - N/2)).
 * x: (N..., K) ComplexF32 device; U: (Nt, K) or NULL for I(Nt); cmaps: ncoils
 * device pointers or NULL (one coil, S = 1).  out: (nsamp, nt, C) ComplexF32;
 * samples removed by the trajectory's mask are written as 0.  Synchronises
 * the stream (its workspace is released on return). */
int msr_forward(msr_ctx *ctx, const msr_traj *traj, int d, const int64_t *img_shape,
                const void *x, int K, const void *U, int u_real, int ncoils,
                const void *const *cmaps, void *out, void *stream);

/* ------------------------------------------------------------------- CG */
/* cg!(x, A, b; maxiter, reltol, abstol) of IterativeSolvers.jl 0.9 (caller of
 * the apply: test/reconstruct_radial_gpu.jl:99, src/CoilMaps.jl:136) run on
 * the device with FP64 inner products, norms and step lengths.  initially_zero
 * = 1 reproduces cg(A, b) (x starts at 0 and is not read).  mode: MSR_CG_*.
 * precision: MSR_PREC_* (MSR_CG_HOST supports MSR_PREC_F32 and
 * MSR_PREC_MIXED).  x and b are ComplexF32 in every precision.
 * iters_out / residual_out (host, may be NULL) receive the iteration count and
 * final residual norm; reading them synchronises the stream. */
int msr_cg(msr_normalop *op, void *x, const void *b, int maxiter, double reltol, double abstol,
           int initially_zero, int mode, int precision, int *iters_out, double *residual_out,
           void *stream);

/* msr_cg in MSR_CG_DEVICE mode that allocates nothing: no hipMalloc, hipFree
 * or allocating library call, so that it can run under HIP graph capture or
 * in loops that must not allocate.  No counterpart in the reference; for
 * callers that cannot allocate device memory.  Same arguments and result as
 * msr_cg (the same kernels in the same order).  The CG vectors, scalars and
 * reduction partials live in one library-owned device array of
 * MSR_CG_NOALLOC_BYTES per device (held from the library's load to the end
 * of the process); the operator's own apply scratch must already be sized
 * for `precision` (it is for MSR_PREC_F32 / MSR_PREC_MIXED from creation, and
 * for MSR_PREC_F64 after any msr_normalop_apply_f64 or msr_cg at that
 * precision).  ArgumentError if the solve needs more than
 * MSR_CG_NOALLOC_BYTES (about 64 bytes per unknown at MSR_PREC_F64, so up to
 * about 1 M unknowns, e.g. 256^2 x K = 16) or the apply scratch is not sized.
 * Every call on one device shares that array (the start kernels rewrite
 * everything a call reads): calls that overlap on one device, from any
 * thread, stream or operator, race and give wrong results. */
#define MSR_CG_NOALLOC_BYTES (64u << 20)
int msr_cg_noalloc(msr_normalop *op, void *x, const void *b, int maxiter, double reltol,
                   double abstol, int initially_zero, int precision, int *iters_out,
                   double *residual_out, void *stream);

/* msr_normalop_apply on ComplexF64 vectors with ComplexF64 arithmetic (the
 * apply of MSR_PREC_F64).  No counterpart in the reference's GPU extension,
 * which is ComplexF32 only; the reference's CPU mul! is generic in the element
 * type (src/NFFTNormalOp.jl:227-267). */
int msr_normalop_apply_f64(msr_normalop *op, void *x, const void *b, double alpha_re,
                           double alpha_im, double beta_re, double beta_im, void *stream);

/* msr_normalop_apply_f64 with the intermediates of MSR_PREC_F64C (the apply
 * msr_cg runs at that precision): ComplexF64 vectors and arithmetic, FFT
 * intermediates stored in 8 bytes per value.  No counterpart in the
 * reference (see msr_normalop_apply_f64). */
int msr_normalop_apply_f64c(msr_normalop *op, void *x, const void *b, double alpha_re,
                            double alpha_im, double beta_re, double beta_im, void *stream);

/* Measured launch configuration of the fused 2D apply at a CG precision's
 * apply mode (MSR_PREC_F32 and MSR_PREC_MIXED share the ComplexF32 apply):
 * the counterpart of the reference's measured FFT planning, plan_fft! with
 * FFTW.MEASURE (src/NFFTNormalOp.jl:71-72,141).  Candidates: pass B coil
 * groups Gb in {1, 2, 4, 8} (up to the coil count) times the two block orders of
 * passes A and C, index 4 log2(Gb) + 2 orderA + orderC; each runs nrep
 * applies (twice over), and must reproduce the first candidate's result bit
 * for bit (MSR_ERR_ASSERTION otherwise).  The fastest is kept for the
 * operator, for operators of the same shape created later on the context,
 * and for batches made afterwards.  times_ms: MSR_TUNE_CONFIGS doubles (ms per
 * apply, NaN for candidates not run) or NULL; *chosen: the kept index, or -1
 * when the operator has no fused 2D passes (3D, Cartesian, line-kernel apply).
 * nrep = 0 measures nothing and reports the current index.  Synchronises the
 * host. */
#define MSR_TUNE_CONFIGS 16
int msr_normalop_tune(msr_normalop *op, int precision, int nrep, double *times_ms, int *chosen,
                      void *stream);

/* The context's tuned choices to / from a text file (the counterpart of
 * FFTW's exported wisdom, which the reference's FFTW.MEASURE planning can
 * reuse): save writes every (shape, apply mode) msr_normalop_tune measured on
 * the context; load adds the file's entries to the context, so operators
 * created afterwards start with them and no tuning runs.  load refuses a file
 * measured on another device type (MSR_ERR_ARGUMENT) and a malformed one
 * (MSR_ERR_IO); *nloaded (may be NULL) = the number of entries. */
int msr_tune_save(msr_ctx *ctx, const char *path);
int msr_tune_load(msr_ctx *ctx, const char *path, int *nloaded);

/* -------------------------------------------------------------- batches */
/* Many reconstructions (uneven image size, coil count and K, 2D and 3D)
 * applied and solved together: each pass of the fused apply is one launch per
 * group of operators that share its FFT length (and K for the kernel
 * multiply), and the CG vector steps are one launch for the whole batch.
 * Every operator's result is the same as from its own msr_normalop_apply /
 * msr_cg.  No counterpart in the reference (which applies one operator per
 * call); it is the batched form of msr_normalop_apply and msr_cg.
 * The batch borrows the operators: they must outlive it, must be distinct,
 * and must not be used by other calls while a batch call runs on them (each
 * operator owns one scratch and one CG state).  Not re-entrant. */
int msr_batch_create(msr_ctx *ctx, int nops, msr_normalop *const *ops, msr_batch **out);
int msr_batch_destroy(msr_batch *batch);
/* x[j] = beta x[j] + alpha A_j b[j] for every operator (x, b: nops device pointers). */
int msr_batch_apply(msr_batch *batch, void *const *x, const void *const *b, double alpha_re,
                    double alpha_im, double beta_re, double beta_im, void *stream);
/* msr_cg for every operator in lockstep; an operator that meets its tolerance
 * stops (its blocks are skipped) while the others go on.  mode:
 * MSR_CG_DEVICE or MSR_CG_GRAPH (the whole solve captured once per pointer
 * set).  iters_out / residual_out: nops entries each, or NULL. */
int msr_batch_cg(msr_batch *batch, void *const *x, const void *const *b, int maxiter,
                 double reltol, double abstol, int initially_zero, int mode, int precision,
                 int *iters_out, double *residual_out, void *stream);
/* The same solve in steps, for schedulers that admit and retire jobs between
 * iterations (continuous batching): start records each operator's CG state on
 * the operator, so a later batch built from any mix of started operators can
 * iterate and finish them.  finish writes the ComplexF32 solutions. */
int msr_batch_cg_start(msr_batch *batch, void *const *x, const void *const *b, int maxiter,
                       double reltol, double abstol, int initially_zero, int precision,
                       void *stream);
int msr_batch_cg_iterate(msr_batch *batch, int niter, void *stream);
int msr_batch_cg_finish(msr_batch *batch, void *stream);
/* Per-operator iteration count, residual norm and done flag (nops entries
 * each, any may be NULL); synchronises the stream. */
int msr_batch_cg_status(msr_batch *batch, int *iters, double *residual, int *done, void *stream);

/* ------------------------------------------------------ shared kernels */
/* A further operator on `proto`'s Toeplitz kernel and kmask (held until the
 * last operator using them is destroyed) with its own coil maps (ncoils
 * device pointers, or 0/NULL for S = 1), apply scratch and CG state, so that
 * several solves of one protocol (trajectory and basis) run at once, e.g.
 * in a batch or a server.  No counterpart in the reference (its operator owns
 * its kernel); callers that serve many reconstructions of one protocol. */
int msr_normalop_create_shared(msr_ctx *ctx, const msr_normalop *proto, int ncoils,
                               const void *const *cmaps, void *stream, msr_normalop **out);
/* Device bytes of an operator's own state for a solve at `precision`
 * (coil maps, apply scratch, CG vectors) and of its (shareable) kernel. */
int msr_normalop_footprint(const msr_normalop *op, int precision, size_t *own_bytes,
                           size_t *kernel_bytes);

/* ------------------------------------------- kernel persistence and cache */
/* No counterpart in the reference, whose callers rebuild the kernel per
 * NFFTNormalOp(img_shape, trj, U) or keep Λ and kmask_indcs themselves for the
 * precomputed-kernel constructor (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:22-70).
 * For servers that see the same protocol (trajectory, sample mask, basis,
 * image size) many times.
 *
 * Kernel file: a 128-byte header (magic "MSRTOEPL", version 1, d, K, P,
 * lambda_real, img_shape[3], nkm, protocol key, payload size and checksum,
 * header checksum), then kmask (nkm Int64, 1-based) and the packed kernel
 * (P, nkm) ComplexF32 or Float32: exactly the two arrays of the
 * precomputed-kernel constructor, in Julia column-major order. */

/* 128-bit key of the kernel msr_normalop_create_from_traj would build from
 * (traj, img_shape, K, U, u_real): a hash of the image size, K, the basis
 * values and kind, the masked trajectory points and per-frame sample counts,
 * and the library version, computed on the device.  Synchronises the stream. */
int msr_kernel_key(msr_ctx *ctx, const msr_traj *traj, int d, const int64_t *img_shape, int K,
                   const void *U, int u_real, void *stream, uint64_t *key);
/* Writes op's kernel and kmask to `path` (through a temporary file renamed
 * into place).  key: 2 words stored in the header, or NULL (0).
 * Synchronises the stream. */
int msr_kernel_save(const msr_normalop *op, const char *path, const uint64_t *key, void *stream);
/* NFFTNormalOp(img_shape, Λ, kmask_indcs; cmaps) from a kernel file: checks
 * the header, sizes and checksums (MSR_ERR_IO on any mismatch) and, if
 * expect_key is not NULL, the stored key (MSR_ERR_ARGUMENT); kmask bounds as
 * msr_normalop_create.  Synchronises the stream. */
int msr_kernel_load(msr_ctx *ctx, const char *path, const uint64_t *expect_key, int ncoils,
                    const void *const *cmaps, void *stream, msr_normalop **out);
/* Header of a kernel file (any output may be NULL; img_shape 3 entries, key
 * 2); verify = 1 also reads the payload and checks its checksum.  Host only:
 * needs no device. */
int msr_kernel_file_info(const char *path, int verify, int *d, int64_t *img_shape, int *K,
                         int *lambda_real, int64_t *nkm, uint64_t *key);

/* Kernels by protocol key: in device memory (least recently used dropped
 * first while the cached kernels exceed budget_bytes; a dropped kernel stays
 * alive while operators use it) and, if directory is not NULL, in kernel
 * files <key>.msrk there (read on a memory miss, written after a build; a
 * corrupt file is rebuilt and replaced).  msr_kernel_cache_get returns a new
 * operator (own coil maps, scratch and CG state; destroy it as usual) on the
 * protocol's kernel and reports where the kernel came from. */
#define MSR_KERNEL_BUILT 0
#define MSR_KERNEL_MEMORY 1
#define MSR_KERNEL_DISK 2
typedef struct msr_kernel_cache msr_kernel_cache;
int msr_kernel_cache_create(msr_ctx *ctx, size_t budget_bytes, const char *directory,
                            msr_kernel_cache **out);
int msr_kernel_cache_destroy(msr_kernel_cache *cache);
/* source (may be NULL): MSR_KERNEL_*; key (may be NULL): the protocol key. */
int msr_kernel_cache_get(msr_kernel_cache *cache, const msr_traj *traj, int d,
                         const int64_t *img_shape, int K, const void *U, int u_real, int ncoils,
                         const void *const *cmaps, void *stream, msr_normalop **out, int *source,
                         uint64_t *key);
/* entries, bytes: cached kernels and their device bytes; counts: 3 entries,
 * gets served per MSR_KERNEL_* source (any may be NULL). */
int msr_kernel_cache_stats(msr_kernel_cache *cache, int64_t *entries, size_t *bytes,
                           int64_t *counts);

/* --------------------------------------------------- Cartesian operator */
/* FFTNormalOp and the Integer-trajectory calculate_backprojection
 * (src/FFTNormalOp.jl; src/BackProjection.jl:104-156), which the reference
 * has on the CPU only.  Samples lie on the N grid: integer index k (1..N,
 * centred k-space) falls in FFT-ordered cell mod(k - N/2 - 1, N).  The
 * kernel is the full (K, K) block per cell (Julia Λ[j, k, q]); the
 * operator's kernel info (msr_normalop_info) reports P = K*K and
 * msr_normalop_export_kernel writes the reference's A.prod!.A.Λ (K, K, Nkm)
 * and kmask_indcs.  The operator works in every msr_normalop_* apply, msr_cg,
 * batch and server call (not msr_normalop_apply_percoil, not kernel files). */

/* Integer trajectory (d, nsamp, nt) Int32 device, e.g. traj_cartesian(...;
 * T=Int32) (src/Trajectories.jl:15-33); mask as msr_traj_create.  Given to
 * msr_backprojection it selects the Cartesian backprojection (no density
 * compensation). */
int msr_traj_create_cartesian(msr_ctx *ctx, int d, int64_t nsamp, int64_t nt, const int32_t *trj,
                              const uint8_t *mask, void *stream, msr_traj **out);
/* FFTNormalOp(img_shape, trj, U; cmaps, sample_mask) (src/FFTNormalOp.jl:28-31,
 * calculate_kernel_cartesian :81-99): the kernel accumulated per cell in FP64
 * over the cell's samples in sample order, rounded to the element type of U
 * (ComplexF32, or Float32 with u_real); kmask = cells with Λ[1,1] != 0
 * (:44-47).  traj from msr_traj_create_cartesian. */
int msr_fftnormalop_create_from_traj(msr_ctx *ctx, const msr_traj *traj, int d,
                                     const int64_t *img_shape, int K, const void *U, int u_real,
                                     int ncoils, const void *const *cmaps, void *stream,
                                     msr_normalop **out);
/* FFTNormalOp(M, U; cmaps) (src/FFTNormalOp.jl:33-36, :101-111): M
 * (img_shape..., Nt) Bool (uint8) device sampling mask in centred k-space
 * (not modified; the reference ifftshifts it in place). */
int msr_fftnormalop_create_from_mask(msr_ctx *ctx, int d, const int64_t *img_shape, int64_t nt,
                                     const uint8_t *M, int K, const void *U, int u_real,
                                     int ncoils, const void *const *cmaps, void *stream,
                                     msr_normalop **out);
/* FFTNormalOp(Λ; cmaps) (src/FFTNormalOp.jl:38-64): lambda (K, K, img_shape...)
 * ComplexF32 or Float32 (lambda_real) device; cells with Λ[1,1] == 0 are
 * dropped, the others kept in full. */
int msr_fftnormalop_create(msr_ctx *ctx, int d, const int64_t *img_shape, int K, int lambda_real,
                           const void *lambda, int ncoils, const void *const *cmaps, void *stream,
                           msr_normalop **out);
/* cartesian = 1 for an FFTNormalOp, 0 for an NFFTNormalOp. */
int msr_normalop_kind(const msr_normalop *op, int *cartesian);

/* ------------------------------------------------ coil-map calibration */
/* The GPU part of calculate_coil_maps (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:1-36):
 * calibration region, coil-wise reconstruction, calibration k-space.  The
 * ESPIRiT step that turns the calibration k-space into coil maps
 * (MRICoilSensitivities.jl, CPU) is not part of this library. */

/* mask_calib and trj_calib (CoilMaps.jl:17-20): scale = img_shape ./
 * calib_size rounded to Float32; trj_out = trj .* scale; mask_out[j] = all
 * (|trj[:, j]| .* scale .< 0.5) && mask[j].  trj, trj_out (d, nsamp, nt)
 * Float32; mask (nsamp, nt) uint8 or NULL (all); mask_out (nsamp, nt) uint8.
 * AssertionError if a calib_size is not 2,3,5-smooth (:14). */
int msr_calib_select(msr_ctx *ctx, int d, int64_t nsamp, int64_t nt, const float *trj,
                     const uint8_t *mask, const int64_t *img_shape, const int64_t *calib_size,
                     float *trj_out, uint8_t *mask_out, void *stream);
/* reconstruct_coilwise(data, trj, img_shape; U, sample_mask, Niter_cg)
 * (CoilMaps.jl:56-77): the K = 1 operator of the first basis column U
 * (Nt values, ComplexF32 or Float32 with u_real; NULL for the default
 * ones(Nt)), the per-coil backprojection, and cg!(x_c, AᴴA, b_c;
 * maxiter = niter, reltol = 0) for each coil, all coils in lockstep on one
 * shared kernel.  precision: MSR_PREC_* of the solves (the reference's
 * arithmetic is MSR_PREC_F32).  data (nsamp, nt, ncoils); out (img_shape...,
 * ncoils) ComplexF32.  Synchronises the stream. */
int msr_reconstruct_coilwise(msr_ctx *ctx, const msr_traj *traj, int d, const int64_t *img_shape,
                             const void *data, int ncoils, const void *U, int u_real, int niter,
                             int precision, void *out, void *stream);
/* kbp = fftshift(fft(fftshift(x, dims)), dims) over the image dimensions of
 * each of ncoils images (CoilMaps.jl:22-25): x, out (calib_size..., ncoils)
 * ComplexF32, not aliased.  Synchronises the stream. */
int msr_calib_kspace(msr_ctx *ctx, int d, const int64_t *calib_size, int ncoils, const void *x,
                     void *out, void *stream);

/* --------------------------------------------------------------- server */
/* Continuous batching of CG solves (the serving loop for streams of
 * reconstructions).  No counterpart in the reference.  A job is an operator
 * with its b and x (ComplexF32 device vectors), maxiter and tolerances; it
 * waits in a FIFO queue until it fits: msr_server_step admits queued jobs while
 * the device state of the jobs in flight (msr_normalop_footprint) stays within
 * `budget_bytes` and there are at most `max_active` of them, starts their CG
 * at `precision`, runs `niter` lockstep CG iterations over every job in
 * flight (msr_batch_cg_iterate) and retires the jobs whose CG has stopped
 * (solution written to x).  A job's result is the one its operator gives in
 * msr_cg alone.  One job per operator at a time; the operators must outlive
 * their jobs.  Not re-entrant. */
#define MSR_JOB_QUEUED 0
#define MSR_JOB_RUNNING 1
#define MSR_JOB_DONE 2
typedef struct msr_server msr_server;
int msr_server_create(msr_ctx *ctx, size_t budget_bytes, int max_active, int precision,
                      msr_server **out);
int msr_server_destroy(msr_server *server);
int msr_server_submit(msr_server *server, msr_normalop *op, void *x, const void *b, int maxiter,
                      double reltol, double abstol, int64_t *job_id);
/* Admission order of queued jobs: MSR_SERVER_FIFO (arrival order, the default)
 * or MSR_SERVER_SHORTEST_FIRST (least estimated work first: maxiter x coils x
 * K x grid cells of the operator; ties in arrival order).  With a small
 * max_active the second approximates shortest-job-first scheduling, which
 * lowers the median latency of a mixed stream at high load. */
#define MSR_SERVER_FIFO 0
#define MSR_SERVER_SHORTEST_FIRST 1
int msr_server_set_policy(msr_server *server, int policy);
/* n_active / n_queued (may be NULL): jobs in flight / waiting after the step;
 * synchronises the stream. */
int msr_server_step(msr_server *server, int niter, int *n_active, int *n_queued, void *stream);
/* state: MSR_JOB_*; times_ms (3 entries, may be NULL): time queued, time in
 * flight, total (host clock, to the step that retired the job). */
int msr_server_poll(msr_server *server, int64_t job_id, int *state, int *iters, double *residual,
                    double *times_ms);
/* Drops a finished job's record. */
int msr_server_forget(msr_server *server, int64_t job_id);

#ifdef __cplusplus
}
#endif

#endif /* MSR_H */
