// extern "C" entry points (include/msr.h): argument checks, exception ->
// status-code translation, thread-local error message.
#include <cstring>
#include <string>

#include "batch.hpp"
#include "normalop.hpp"
#include "server.hpp"

namespace msr {
// traj.hip
msr_traj* traj_create(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const float* trj,
                      const uint8_t* mask, hipStream_t st);
// kernel_build.hip
void op_build_from_traj(msr_normalop* op, const msr_traj* traj, const void* U, int u_real,
                        hipStream_t st);
// backproj.hip
void backprojection(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N, const void* data,
                    int ncoils, int K, const void* U, int u_real, const void* const* cmaps,
                    int dcf_mode, const float* dcf, void* out, hipStream_t st, bool accumulate = false);
void backprojection_channelwise(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N,
                                const void* data, int ncoils, int K, const void* U, int u_real,
                                const void* const* cmaps, int dcf_mode, const float* dcf, void* out,
                                hipStream_t st);

// forward.hip
void forward_model(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N, const void* x,
                   int K, const void* U, int u_real, int ncoils, const void* const* cmaps, void* out,
                   hipStream_t st);

// server.hip
msr_server* server_create(msr_ctx* ctx, size_t budget, int max_active, int prec);
int64_t server_submit(msr_server* s, msr_normalop* op, float2* x, const float2* b, int maxiter,
                      double reltol, double abstol);
void server_step(msr_server* s, int niter, int* n_active, int* n_queued, hipStream_t st);
void server_poll(msr_server* s, int64_t id, int* state, int* iters, double* res, double* times);
void server_forget(msr_server* s, int64_t id);
void server_set_policy(msr_server* s, int policy);

// kcache.hip
void kernel_key(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N, int K,
                const void* U, int u_real, hipStream_t st, uint64_t key[2]);
void kernel_save(const msr_normalop* op, const char* path, const uint64_t* key, hipStream_t st);
msr_normalop* kernel_load(msr_ctx* ctx, const char* path, const uint64_t* expect_key,
                          int ncoils, const void* const* cmaps, hipStream_t st);
void kernel_file_info(const char* path, int verify, int* d, int64_t* N, int* K, int* lam_real,
                      int64_t* nkm, uint64_t* key);
msr_kernel_cache* kernel_cache_create(msr_ctx* ctx, size_t budget, const char* dir);
msr_normalop* kernel_cache_get(msr_kernel_cache* c, const msr_traj* traj, int d, const int64_t* N,
                               int K, const void* U, int u_real, int ncoils,
                               const void* const* cmaps, hipStream_t st, int* source,
                               uint64_t* key_out);
void kernel_cache_stats(msr_kernel_cache* c, int64_t* entries, size_t* bytes, int64_t* counts);
void kernel_cache_destroy(msr_kernel_cache* c);
msr_ctx* cache_ctx(const msr_kernel_cache* c);
// cartesian.hip
msr_traj* cart_traj_create(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const int32_t* trj,
                           const uint8_t* mask, hipStream_t st);
void cart_build_kernel(msr_normalop* op, const msr_traj* traj, const uint8_t* M, int64_t nt_mask,
                       const void* U, int u_real, hipStream_t st);
void cart_load_dense(msr_normalop* op, const void* lambda, hipStream_t st);
// coilmaps.hip
void calib_select(int d, int64_t nsamp, int64_t nt, const float* trj, const uint8_t* mask,
                  const int64_t* N, const int64_t* Nc, float* trj_out, uint8_t* mask_out,
                  hipStream_t st);
void calib_kspace(msr_ctx* ctx, int d, const int64_t* Nc, int ncoils, const float2* x,
                  float2* out, hipStream_t st);
void reconstruct_coilwise(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N,
                          const void* data, int ncoils, const void* U, int u_real, int niter,
                          int prec, float2* out, hipStream_t st);

static thread_local std::string g_err;
void set_last_error(const std::string& s) { g_err = s; }
}  // namespace msr

using namespace msr;

template <class F>
static int guarded(F&& f) {
    try {
        f();
        return MSR_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return MSR_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MSR_ERR_ERROR;
    } catch (...) {
        set_last_error("unknown error");
        return MSR_ERR_ERROR;
    }
}

#define MSR_NOTNULL(p) MSR_REQUIRE((p) != nullptr, MSR_ERR_ARGUMENT, "%s is NULL", #p)

extern "C" {

size_t msr_last_error(char* buf, size_t len) {
    if (buf && len) {
        size_t n = std::min(len - 1, g_err.size());
        memcpy(buf, g_err.data(), n);
        buf[n] = 0;
    }
    return g_err.size();
}

int msr_version(void) { return 10000; }

int msr_ctx_create(int device, msr_ctx** out) {
    return guarded([&] {
        MSR_NOTNULL(out);
        int n = 0;
        MSR_HIP(hipGetDeviceCount(&n));
        MSR_REQUIRE(device >= 0 && device < n, MSR_ERR_ARGUMENT, "device %d of %d", device, n);
        auto* c = new msr_ctx;
        c->device = device;
        try {
            ensure_device(c);
            cg_noalloc_setup(device);  // device address of msr_cg_noalloc's scratch array
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

int msr_ctx_destroy(msr_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return;
        ensure_device(ctx);
        delete ctx;
    });
}

int msr_traj_create(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const float* trj,
                    const uint8_t* mask, void* stream, msr_traj** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        *out = traj_create(ctx, d, nsamp, nt, trj, mask, (hipStream_t)stream);
    });
}

int msr_traj_info(const msr_traj* traj, int64_t* n_masked, int64_t* nt) {
    return guarded([&] {
        MSR_NOTNULL(traj);
        if (n_masked) *n_masked = traj->M;
        if (nt) *nt = traj->nt;
    });
}

int msr_traj_destroy(msr_traj* traj) {
    return guarded([&] {
        if (!traj) return;
        ensure_device(traj->ctx);
        delete traj;
    });
}

int msr_normalop_create_from_traj(msr_ctx* ctx, const msr_traj* traj, int d,
                                  const int64_t* img_shape, int K, const void* U, int u_real,
                                  int ncoils, const void* const* cmaps, void* stream,
                                  msr_normalop** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(U);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        MSR_REQUIRE(traj->d == d, MSR_ERR_ARGUMENT,
                    "trajectory has %d dimensions, img_shape %d", traj->d, d);
        hipStream_t st = (hipStream_t)stream;
        std::unique_ptr<msr_normalop> op(op_new(ctx, d, img_shape, K, u_real ? 1 : 0));
        op_set_cmaps(op.get(), ncoils, cmaps, st);
        op_build_from_traj(op.get(), traj, U, u_real, st);
        op_alloc_scratch(op.get());
        MSR_HIP(hipStreamSynchronize(st));
        *out = op.release();
    });
}

int msr_normalop_create(msr_ctx* ctx, int d, const int64_t* img_shape, int64_t P,
                        int lambda_real, const void* lambda, const int64_t* kmask, int64_t nkm,
                        int ncoils, const void* const* cmaps, void* stream, msr_normalop** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(out);
        MSR_REQUIRE(nkm >= 0, MSR_ERR_ARGUMENT, "nkm = %lld", (long long)nkm);
        MSR_REQUIRE(nkm == 0 || (lambda && kmask), MSR_ERR_ARGUMENT, "lambda/kmask is NULL");
        ensure_device(ctx);
        // Ncoeff = (isqrt(8P+1) - 1) ÷ 2 (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:33)
        int64_t s = (int64_t)std::llround(std::sqrt((double)(8 * P + 1)));
        while (s * s > 8 * P + 1) --s;
        while ((s + 1) * (s + 1) <= 8 * P + 1) ++s;
        const int K = (int)((s - 1) / 2);
        MSR_REQUIRE(K >= 1 && (int64_t)K * (K + 1) / 2 == P, MSR_ERR_ARGUMENT,
                    "packed kernel axis P = %lld is not K(K+1)/2", (long long)P);
        hipStream_t st = (hipStream_t)stream;
        std::unique_ptr<msr_normalop> op(op_new(ctx, d, img_shape, K, lambda_real ? 1 : 0));
        op_set_cmaps(op.get(), ncoils, cmaps, st);
        op_load_packed(op.get(), lambda, kmask, nkm, st);
        op_alloc_scratch(op.get());
        MSR_HIP(hipStreamSynchronize(st));
        *out = op.release();
    });
}

int msr_normalop_info(const msr_normalop* op, int* K, int64_t* P, int64_t* nkm, int* lambda_real) {
    return guarded([&] {
        MSR_NOTNULL(op);
        if (K) *K = op->K;
        if (P) *P = op->P;
        if (nkm) *nkm = op->nkm;
        if (lambda_real) *lambda_real = op->lam_real;
    });
}

int msr_normalop_export_kernel(const msr_normalop* op, void* lambda_out, int64_t* kmask_out,
                               void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        ensure_device(op->ctx);
        op_export(op, lambda_out, kmask_out, (hipStream_t)stream);
    });
}

int msr_normalop_apply(msr_normalop* op, void* x, const void* b, double alpha_re,
                       double alpha_im, double beta_re, double beta_im, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_NOTNULL(x);
        MSR_NOTNULL(b);
        MSR_REQUIRE(x != b, MSR_ERR_ARGUMENT, "x and b must not alias");
        ensure_device(op->ctx);
        op_apply(op, (float2*)x, (const float2*)b, make_float2((float)alpha_re, (float)alpha_im),
                 make_float2((float)beta_re, (float)beta_im), beta_re == 0.0 && beta_im == 0.0,
                 (hipStream_t)stream);
    });
}

int msr_normalop_apply_percoil(msr_normalop* op, void* x, const void* b, double alpha_re,
                               double alpha_im, double beta_re, double beta_im, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_NOTNULL(x);
        MSR_NOTNULL(b);
        MSR_REQUIRE(x != b, MSR_ERR_ARGUMENT, "x and b must not alias");
        ensure_device(op->ctx);
        op_apply_percoil(op, (float2*)x, (const float2*)b,
                         make_float2((float)alpha_re, (float)alpha_im),
                         make_float2((float)beta_re, (float)beta_im),
                         beta_re == 0.0 && beta_im == 0.0, (hipStream_t)stream);
    });
}

int msr_normalop_destroy(msr_normalop* op) {
    return guarded([&] {
        if (!op) return;
        ensure_device(op->ctx);
        delete op;
    });
}

int msr_backprojection(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* img_shape,
                       const void* data, int ncoils, int K, const void* U, int u_real,
                       const void* const* cmaps, int dcf_mode, const float* dcf, void* out,
                       void* stream) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(data);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        backprojection(ctx, traj, d, img_shape, data, ncoils, K, U, u_real, cmaps, dcf_mode, dcf,
                       out, (hipStream_t)stream);
    });
}

int msr_backprojection_channelwise(msr_ctx* ctx, const msr_traj* traj, int d,
                                   const int64_t* img_shape, const void* data, int ncoils, int K,
                                   const void* U, int u_real, const void* const* cmaps, int dcf_mode,
                                   const float* dcf, void* out, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(data);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        backprojection_channelwise(ctx, traj, d, img_shape, data, ncoils, K, U, u_real, cmaps,
                                   dcf_mode, dcf, out, (hipStream_t)stream);
    });
}

int msr_forward(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* img_shape,
                const void* x, int K, const void* U, int u_real, int ncoils,
                const void* const* cmaps, void* out, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(x);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        forward_model(ctx, traj, d, img_shape, x, K, U, u_real, cmaps ? ncoils : 0, cmaps, out,
                      (hipStream_t)stream);
    });
}

int msr_normalop_apply_f64(msr_normalop* op, void* x, const void* b, double alpha_re,
                           double alpha_im, double beta_re, double beta_im, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_NOTNULL(x);
        MSR_NOTNULL(b);
        MSR_REQUIRE(x != b, MSR_ERR_ARGUMENT, "x and b must not alias");
        ensure_device(op->ctx);
        op_apply_f64(op, (double2*)x, (const double2*)b, make_double2(alpha_re, alpha_im),
                     make_double2(beta_re, beta_im), beta_re == 0.0 && beta_im == 0.0,
                     (hipStream_t)stream);
    });
}

int msr_normalop_apply_f64c(msr_normalop* op, void* x, const void* b, double alpha_re,
                            double alpha_im, double beta_re, double beta_im, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_NOTNULL(x);
        MSR_NOTNULL(b);
        MSR_REQUIRE(x != b, MSR_ERR_ARGUMENT, "x and b must not alias");
        ensure_device(op->ctx);
        op_apply_f64(op, (double2*)x, (const double2*)b, make_double2(alpha_re, alpha_im),
                     make_double2(beta_re, beta_im), beta_re == 0.0 && beta_im == 0.0,
                     (hipStream_t)stream, AM_F64C);
    });
}

int msr_normalop_tune(msr_normalop* op, int precision, int nrep, double* times_ms, int* chosen,
                      void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_REQUIRE(prec_valid(precision), MSR_ERR_ARGUMENT, "precision = %d", precision);
        static_assert(TUNE_CONFIGS == MSR_TUNE_CONFIGS, "tuner candidate count");
        ensure_device(op->ctx);
        op_tune(op, apply_mode(precision), nrep, times_ms, chosen, (hipStream_t)stream);
    });
}

int msr_tune_save(msr_ctx* ctx, const char* path) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(path);
        tune_save(ctx, path);
    });
}

int msr_tune_load(msr_ctx* ctx, const char* path, int* nloaded) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(path);
        const int n = tune_load(ctx, path);
        if (nloaded) *nloaded = n;
    });
}

int msr_cg(msr_normalop* op, void* x, const void* b, int maxiter, double reltol, double abstol,
           int initially_zero, int mode, int precision, int* iters_out, double* residual_out,
           void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_NOTNULL(x);
        MSR_NOTNULL(b);
        MSR_REQUIRE(x != b, MSR_ERR_ARGUMENT, "x and b must not alias");
        MSR_REQUIRE(maxiter >= 0, MSR_ERR_ARGUMENT, "maxiter = %d", maxiter);
        ensure_device(op->ctx);
        op_cg(op, (float2*)x, (const float2*)b, maxiter, reltol, abstol, initially_zero != 0, mode,
              precision, iters_out, residual_out, (hipStream_t)stream);
    });
}

int msr_cg_noalloc(msr_normalop* op, void* x, const void* b, int maxiter, double reltol,
                   double abstol, int initially_zero, int precision, int* iters_out,
                   double* residual_out, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_NOTNULL(x);
        MSR_NOTNULL(b);
        MSR_REQUIRE(x != b, MSR_ERR_ARGUMENT, "x and b must not alias");
        ensure_device(op->ctx);
        op_cg_noalloc(op, (float2*)x, (const float2*)b, maxiter, reltol, abstol,
                      initially_zero != 0, precision, iters_out, residual_out,
                      (hipStream_t)stream);
    });
}

int msr_batch_create(msr_ctx* ctx, int nops, msr_normalop* const* ops, msr_batch** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        *out = batch_create(ctx, nops, ops);
    });
}

int msr_batch_destroy(msr_batch* batch) {
    return guarded([&] {
        if (!batch) return;
        ensure_device(batch->ctx);
        delete batch;
    });
}

int msr_batch_apply(msr_batch* batch, void* const* x, const void* const* b, double alpha_re,
                    double alpha_im, double beta_re, double beta_im, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(batch);
        ensure_device(batch->ctx);
        batch_apply(batch, (float2* const*)x, (const float2* const*)b,
                    make_float2((float)alpha_re, (float)alpha_im),
                    make_float2((float)beta_re, (float)beta_im), beta_re == 0.0 && beta_im == 0.0,
                    (hipStream_t)stream);
    });
}

int msr_batch_cg(msr_batch* batch, void* const* x, const void* const* b, int maxiter,
                 double reltol, double abstol, int initially_zero, int mode, int precision,
                 int* iters_out, double* residual_out, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(batch);
        ensure_device(batch->ctx);
        batch_cg(batch, (float2* const*)x, (const float2* const*)b, maxiter, reltol, abstol,
                 initially_zero != 0, mode, precision, (hipStream_t)stream);
        if (iters_out || residual_out)
            batch_cg_status(batch, iters_out, residual_out, nullptr, (hipStream_t)stream);
    });
}

int msr_batch_cg_start(msr_batch* batch, void* const* x, const void* const* b, int maxiter,
                       double reltol, double abstol, int initially_zero, int precision,
                       void* stream) {
    return guarded([&] {
        MSR_NOTNULL(batch);
        ensure_device(batch->ctx);
        batch_cg_start(batch, (float2* const*)x, (const float2* const*)b, maxiter, reltol, abstol,
                       initially_zero != 0, precision, (hipStream_t)stream);
    });
}

int msr_batch_cg_iterate(msr_batch* batch, int niter, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(batch);
        ensure_device(batch->ctx);
        batch_cg_iterate(batch, niter, (hipStream_t)stream);
    });
}

int msr_batch_cg_finish(msr_batch* batch, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(batch);
        ensure_device(batch->ctx);
        batch_cg_finish(batch, (hipStream_t)stream);
    });
}

int msr_batch_cg_status(msr_batch* batch, int* iters, double* residual, int* done, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(batch);
        ensure_device(batch->ctx);
        batch_cg_status(batch, iters, residual, done, (hipStream_t)stream);
    });
}

}  // extern "C"

int msr_normalop_create_shared(msr_ctx* ctx, const msr_normalop* proto, int ncoils,
                               const void* const* cmaps, void* stream, msr_normalop** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(proto);
        MSR_NOTNULL(out);
        MSR_REQUIRE(proto->ctx == ctx, MSR_ERR_ARGUMENT, "the prototype belongs to another context");
        ensure_device(ctx);
        hipStream_t st = (hipStream_t)stream;
        *out = op_on_kernel(ctx, proto->d, proto->N, proto->K, proto->lam_real, proto->nkm,
                            proto->lam, proto->kmask, ncoils, cmaps, st, proto->cart);
    });
}

int msr_normalop_footprint(const msr_normalop* op, int precision, size_t* own_bytes,
                           size_t* kernel_bytes) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_REQUIRE(prec_valid(precision), MSR_ERR_ARGUMENT, "precision = %d", precision);
        if (own_bytes) *own_bytes = op_footprint(op, precision);
        if (kernel_bytes) *kernel_bytes = op_kernel_bytes(op);
    });
}

int msr_server_create(msr_ctx* ctx, size_t budget_bytes, int max_active, int precision,
                      msr_server** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(out);
        *out = server_create(ctx, budget_bytes, max_active, precision);
    });
}

int msr_server_destroy(msr_server* server) {
    return guarded([&] {
        if (!server) return;
        ensure_device(server->ctx);
        delete server;
    });
}

int msr_server_submit(msr_server* server, msr_normalop* op, void* x, const void* b, int maxiter,
                      double reltol, double abstol, int64_t* job_id) {
    return guarded([&] {
        MSR_NOTNULL(server);
        MSR_NOTNULL(op);
        MSR_NOTNULL(x);
        MSR_NOTNULL(b);
        const int64_t id = server_submit(server, op, (float2*)x, (const float2*)b, maxiter, reltol,
                                         abstol);
        if (job_id) *job_id = id;
    });
}

int msr_server_step(msr_server* server, int niter, int* n_active, int* n_queued, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(server);
        ensure_device(server->ctx);
        server_step(server, niter, n_active, n_queued, (hipStream_t)stream);
    });
}

int msr_server_poll(msr_server* server, int64_t job_id, int* state, int* iters, double* residual,
                    double* times_ms) {
    return guarded([&] {
        MSR_NOTNULL(server);
        server_poll(server, job_id, state, iters, residual, times_ms);
    });
}

int msr_server_forget(msr_server* server, int64_t job_id) {
    return guarded([&] {
        MSR_NOTNULL(server);
        server_forget(server, job_id);
    });
}

// ------------------------------------------------- kernel persistence and cache
int msr_kernel_key(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* img_shape, int K,
                   const void* U, int u_real, void* stream, uint64_t* key) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(U);
        MSR_NOTNULL(key);
        ensure_device(ctx);
        kernel_key(ctx, traj, d, img_shape, K, U, u_real, (hipStream_t)stream, key);
    });
}

int msr_kernel_save(const msr_normalop* op, const char* path, const uint64_t* key, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(op);
        MSR_NOTNULL(path);
        ensure_device(op->ctx);
        kernel_save(op, path, key, (hipStream_t)stream);
    });
}

int msr_kernel_load(msr_ctx* ctx, const char* path, const uint64_t* expect_key, int ncoils,
                    const void* const* cmaps, void* stream, msr_normalop** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(path);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        *out = kernel_load(ctx, path, expect_key, ncoils, cmaps, (hipStream_t)stream);
    });
}

int msr_kernel_file_info(const char* path, int verify, int* d, int64_t* img_shape, int* K,
                         int* lambda_real, int64_t* nkm, uint64_t* key) {
    return guarded([&] {
        MSR_NOTNULL(path);
        kernel_file_info(path, verify, d, img_shape, K, lambda_real, nkm, key);
    });
}

int msr_kernel_cache_create(msr_ctx* ctx, size_t budget_bytes, const char* directory,
                            msr_kernel_cache** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(out);
        *out = kernel_cache_create(ctx, budget_bytes, directory);
    });
}

int msr_kernel_cache_destroy(msr_kernel_cache* cache) {
    return guarded([&] {
        if (cache) kernel_cache_destroy(cache);
    });
}

int msr_kernel_cache_get(msr_kernel_cache* cache, const msr_traj* traj, int d,
                         const int64_t* img_shape, int K, const void* U, int u_real, int ncoils,
                         const void* const* cmaps, void* stream, msr_normalop** out, int* source,
                         uint64_t* key) {
    return guarded([&] {
        MSR_NOTNULL(cache);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(U);
        MSR_NOTNULL(out);
        ensure_device(traj->ctx);
        MSR_REQUIRE(traj->ctx == cache_ctx(cache), MSR_ERR_ARGUMENT,
                    "the trajectory belongs to another context");
        *out = kernel_cache_get(cache, traj, d, img_shape, K, U, u_real, ncoils, cmaps,
                                (hipStream_t)stream, source, key);
    });
}

int msr_kernel_cache_stats(msr_kernel_cache* cache, int64_t* entries, size_t* bytes,
                           int64_t* counts) {
    return guarded([&] {
        MSR_NOTNULL(cache);
        kernel_cache_stats(cache, entries, bytes, counts);
    });
}

// ------------------------------------------------------ coil-map calibration
int msr_calib_select(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const float* trj,
                     const uint8_t* mask, const int64_t* img_shape, const int64_t* calib_size,
                     float* trj_out, uint8_t* mask_out, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(trj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(calib_size);
        MSR_NOTNULL(trj_out);
        MSR_NOTNULL(mask_out);
        ensure_device(ctx);
        calib_select(d, nsamp, nt, trj, mask, img_shape, calib_size, trj_out, mask_out,
                     (hipStream_t)stream);
    });
}

int msr_reconstruct_coilwise(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* img_shape,
                             const void* data, int ncoils, const void* U, int u_real, int niter,
                             int precision, void* out, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(data);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        reconstruct_coilwise(ctx, traj, d, img_shape, data, ncoils, U, u_real, niter, precision,
                             (float2*)out, (hipStream_t)stream);
    });
}

int msr_calib_kspace(msr_ctx* ctx, int d, const int64_t* calib_size, int ncoils, const void* x,
                     void* out, void* stream) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(calib_size);
        MSR_NOTNULL(x);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        calib_kspace(ctx, d, calib_size, ncoils, (const float2*)x, (float2*)out,
                     (hipStream_t)stream);
    });
}

// ------------------------------------------------------ Cartesian operator
int msr_traj_create_cartesian(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const int32_t* trj,
                              const uint8_t* mask, void* stream, msr_traj** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        *out = cart_traj_create(ctx, d, nsamp, nt, trj, mask, (hipStream_t)stream);
    });
}

int msr_fftnormalop_create_from_traj(msr_ctx* ctx, const msr_traj* traj, int d,
                                     const int64_t* img_shape, int K, const void* U, int u_real,
                                     int ncoils, const void* const* cmaps, void* stream,
                                     msr_normalop** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(traj);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(U);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        hipStream_t st = (hipStream_t)stream;
        std::unique_ptr<msr_normalop> op(op_new(ctx, d, img_shape, K, u_real ? 1 : 0, 1));
        op_set_cmaps(op.get(), ncoils, cmaps, st);
        cart_build_kernel(op.get(), traj, nullptr, 0, U, u_real, st);
        op_alloc_scratch(op.get());
        MSR_HIP(hipStreamSynchronize(st));
        *out = op.release();
    });
}

int msr_fftnormalop_create_from_mask(msr_ctx* ctx, int d, const int64_t* img_shape, int64_t nt,
                                     const uint8_t* M, int K, const void* U, int u_real,
                                     int ncoils, const void* const* cmaps, void* stream,
                                     msr_normalop** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(M);
        MSR_NOTNULL(U);
        MSR_NOTNULL(out);
        MSR_REQUIRE(nt >= 1, MSR_ERR_ARGUMENT, "Nt = %lld", (long long)nt);
        ensure_device(ctx);
        hipStream_t st = (hipStream_t)stream;
        std::unique_ptr<msr_normalop> op(op_new(ctx, d, img_shape, K, u_real ? 1 : 0, 1));
        op_set_cmaps(op.get(), ncoils, cmaps, st);
        cart_build_kernel(op.get(), nullptr, M, nt, U, u_real, st);
        op_alloc_scratch(op.get());
        MSR_HIP(hipStreamSynchronize(st));
        *out = op.release();
    });
}

int msr_fftnormalop_create(msr_ctx* ctx, int d, const int64_t* img_shape, int K, int lambda_real,
                           const void* lambda, int ncoils, const void* const* cmaps, void* stream,
                           msr_normalop** out) {
    return guarded([&] {
        MSR_NOTNULL(ctx);
        MSR_NOTNULL(img_shape);
        MSR_NOTNULL(lambda);
        MSR_NOTNULL(out);
        ensure_device(ctx);
        hipStream_t st = (hipStream_t)stream;
        std::unique_ptr<msr_normalop> op(op_new(ctx, d, img_shape, K, lambda_real ? 1 : 0, 1));
        op_set_cmaps(op.get(), ncoils, cmaps, st);
        cart_load_dense(op.get(), lambda, st);
        op_alloc_scratch(op.get());
        MSR_HIP(hipStreamSynchronize(st));
        *out = op.release();
    });
}

int msr_normalop_kind(const msr_normalop* op, int* cartesian) {
    return guarded([&] {
        MSR_NOTNULL(op);
        if (cartesian) *cartesian = op->cart;
    });
}

int msr_server_set_policy(msr_server* server, int policy) {
    return guarded([&] {
        MSR_NOTNULL(server);
        server_set_policy(server, policy);
    });
}
