// Normal-operator object: construction from a packed kernel (the reference's
// NFFTNormalOp(img_shape, Λ, kmask_indcs; cmaps),
// ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:22-70), dense device layout of
// the kernel, fused and per-coil applies, kernel export, device CG.
//
// Dense kernel layout (2D): lam[(qx * P + p) * 2Ny + qy] for every cell of the
// 2Nx x 2Ny grid, zero off the kmask, so that pass B reads one contiguous
// (P, 2Ny) slab per grid column.  For 3D see apply3d.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>

#include "apply3d.hpp"
#include "fft.hpp"
#include "normalop.hpp"

namespace msr {

// ------------------------------------------------------------------ kernels
__global__ void fill_kernel(float* __restrict__ p, int64_t n, float v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_floats(void* p, int64_t n, float v, hipStream_t st) {
    if (n <= 0) return;
    int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, st, (float*)p, n, v);
    MSR_CHECK_LAUNCH();
}

// linear 0-based grid index q (x fastest) -> dense kernel offset of entry p
// (cell_major: the Cartesian kernel, lam[q * P + p])
__device__ __forceinline__ int64_t dense_offset(int d, const int64_t* n2, int P, int64_t q, int p,
                                                int cell_major) {
    if (cell_major) return q * P + p;
    if (d == 2) {
        const int64_t qx = q % n2[0], qy = q / n2[0];
        return ((qx * P) + p) * n2[1] + qy;
    }
    // 3D: lam[((qy * n2x + qx) * P + p) * n2z + qz]
    const int64_t qx = q % n2[0];
    const int64_t r = q / n2[0];
    const int64_t qy = r % n2[1], qz = r / n2[1];
    return (((qy * n2[0] + qx) * P) + p) * n2[2] + qz;
}

struct GridDims {
    int64_t n[3];
    int cell_major = 0;
};

template <class LT>
__global__ void scatter_kernel(const LT* __restrict__ packed, const int64_t* __restrict__ kmask,
                               int64_t nkm, int P, int d, GridDims g, LT* __restrict__ dense,
                               int* __restrict__ bad) {
    const int64_t G = g.n[0] * g.n[1] * (d == 3 ? g.n[2] : 1);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nkm * P;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / P;
        const int p = (int)(e - i * P);
        const int64_t q1 = kmask[i];
        if (q1 < 1 || q1 > G) {
            *bad = 1;
            continue;
        }
        dense[dense_offset(d, g.n, P, q1 - 1, p, g.cell_major)] = packed[e];
    }
}

template <class LT>
__global__ void gather_kernel(const LT* __restrict__ dense, const int64_t* __restrict__ kmask,
                              int64_t nkm, int P, int d, GridDims g, LT* __restrict__ packed) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nkm * P;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / P;
        const int p = (int)(e - i * P);
        packed[e] = dense[dense_offset(d, g.n, P, kmask[i] - 1, p, g.cell_major)];
    }
}

// ------------------------------------------------------------------ helpers
static int env_int(const char* name, int dflt) {
    const char* s = getenv(name);
    return s ? atoi(s) : dflt;
}

static std::array<int64_t, 7> tune_key(const msr_normalop* op, int am) {
    return {op->N[0], op->N[1], op->K, op->C, op->lam_real, op->identity_cmaps ? 1 : 0, am};
}

bool tuned_knobs(const msr_normalop* op, int am, ApplyKnobs* out) {
    std::lock_guard<std::mutex> guard(op->ctx->lock);
    auto it = op->ctx->tuned.find(tune_key(op, am));
    if (it == op->ctx->tuned.end()) return false;
    *out = it->second;
    return true;
}

void remember_knobs(const msr_normalop* op, int am, const ApplyKnobs& k) {
    std::lock_guard<std::mutex> guard(op->ctx->lock);
    op->ctx->tuned[tune_key(op, am)] = k;
}

void op_alloc_scratch(msr_normalop* op) {
    const int64_t ny = op->N[1];
    op->Cb = 1;
    if (!op->cart && !op->fused) {
        op->G = 1;
        op->Gb = 1;
        op->Cb = lines_chunk_coils(op);
    } else if (op->d == 2 && !op->cart) {
        const int tiles = tiles2d((int)ny, 8);
        int G = (int)ceil_div(512, (int64_t)tiles * op->K);
        G = std::max(1, std::min(G, op->C));
        op->G = env_int("MSR_G", G);
        op->G = std::max(1, std::min(op->G, op->C));
        const int64_t colblocks = 2 * op->N[0];
        int Gb = colblocks >= 1024 ? 1 : 2;
        op->Gb = std::max(1, std::min(env_int("MSR_GB", Gb), op->C));
        // untuned knobs: the orders measured fastest at 256^2 / 20 coils / K = 5
        // (pass A tile-fastest only with 16-byte intermediates); a tuned choice
        // for the same shape on this context replaces them
        for (int am = 0; am < 3; ++am) {
            op->knobs[am].Gb = op->Gb;
            op->knobs[am].order_a = am == AM_F64 ? 0 : 1;
            op->knobs[am].order_c = 1;
            ApplyKnobs t;
            if (tuned_knobs(op, am, &t)) op->knobs[am] = t;
        }
    } else {
        op->G = 1;
        op->Gb = 1;
    }
    op_ensure_scratch(op, false);
}

// Intermediates for the apply precision (grow-only: ComplexF64 needs twice the
// bytes).  Re-allocation changes the addresses, which the graph keys include.
void op_ensure_scratch(msr_normalop* op, bool f64) {
    // the twiddle tables of the apply are device allocations too: create them
    // here, outside any stream capture
    const int pad = op->cart ? 1 : 2;
    for (int a = 0; a < op->d; ++a) {
        if (f64) (void)twiddle_table64(op->ctx, (int)(pad * op->N[a]));
        else (void)twiddle_table(op->ctx, (int)(pad * op->N[a]));
    }
    const size_t es = f64 ? sizeof(double2) : sizeof(float2);
    if (op->scratch_es >= es) return;
    const int64_t nx = op->N[0], ny = op->N[1], nz = op->N[2];
    if (op->cart) {  // two ping-pong grids (C, K, N...)
        op->T.alloc(es * (size_t)op->C * op->K * op->V);
        op->T2.alloc(es * (size_t)op->C * op->K * op->V);
    } else if (!op->fused) {  // two ping-pong grids (Cb, K, 2N...)
        op->T.alloc(lines_scratch_bytes(op, es));
        op->T2.alloc(lines_scratch_bytes(op, es));
    } else if (op->d == 2) {
        op->T.alloc(es * (size_t)op->C * op->K * (2 * nx) * ny);
    } else {
        op->T.alloc(es * (size_t)op->C * op->K * nz * ny * 2 * nx);
        op->T2.alloc(es * (size_t)op->C * op->K * nz * 2 * ny * 2 * nx);
    }
    op->partial.alloc(es * (size_t)op->G * op->K * op->V);
    op->scratch_es = es;
}

static void check_shape_cart(int d, const int64_t* N) {
    MSR_REQUIRE(d == 2 || d == 3, MSR_ERR_ARGUMENT, "image dimension d = %d, expected 2 or 3", d);
    for (int i = 0; i < d; ++i)
        MSR_REQUIRE(N[i] >= 1 && fft_len_supported(N[i]), MSR_ERR_UNSUPPORTED,
                    "Cartesian grid size N[%d] = %lld has no FFT kernel in this build", i,
                    (long long)N[i]);
}

// The fused passes take 2N in the fused length set up to 512 along the axes
// they transform in LDS (x and y in 2D, x and z in 3D) and K <= 8; every other
// shape the line kernels cover runs the line-kernel apply.
static bool fused_shape(int d, const int64_t* N, int K) {
    if (K > 8) return false;
    for (int i = 0; i < d; ++i) {
        if (d == 3 && i == 1) continue;  // the 3D y pass runs on the line kernels
        if (!fused_len_supported(2 * N[i]) || 2 * N[i] > 512) return false;
    }
    return true;
}

static void check_shape(int d, const int64_t* N) {
    MSR_REQUIRE(d == 2 || d == 3, MSR_ERR_ARGUMENT, "image dimension d = %d, expected 2 or 3", d);
    for (int i = 0; i < d; ++i) {
        MSR_REQUIRE(N[i] >= 1, MSR_ERR_ARGUMENT, "img_shape[%d] = %lld", i, (long long)N[i]);
        int64_t m = 2 * N[i];
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        MSR_REQUIRE(m == 1, MSR_ERR_ASSERTION,
                    "img_shape_os has to be composed of the prime factors 2, 3, and 5 "
                    "(2N[%d] = %lld)", i, (long long)(2 * N[i]));
        MSR_REQUIRE(fft_len_supported(2 * N[i]), MSR_ERR_UNSUPPORTED,
                    "Toeplitz grid size 2N[%d] = %lld has no FFT kernel in this build "
                    "(2,3,5-smooth lengths up to 2048)", i, (long long)(2 * N[i]));
    }
}

void op_set_cmaps(msr_normalop* op, int ncoils, const void* const* cmaps, hipStream_t st) {
    MSR_REQUIRE(ncoils >= 0, MSR_ERR_ARGUMENT, "ncoils = %d", ncoils);
    if (ncoils == 0 || cmaps == nullptr) {
        op->C = 1;
        op->identity_cmaps = true;
        return;
    }
    op->C = ncoils;
    op->identity_cmaps = false;
    op->cmaps.alloc(sizeof(float2) * (size_t)ncoils * op->V);
    for (int c = 0; c < ncoils; ++c) {
        MSR_REQUIRE(cmaps[c] != nullptr, MSR_ERR_ARGUMENT, "cmaps[%d] is NULL", c);
        MSR_HIP(hipMemcpyAsync(op->cmaps.as<float2>() + (size_t)c * op->V, cmaps[c],
                               sizeof(float2) * op->V, hipMemcpyDeviceToDevice, st));
    }
}

GridDims op_grid(const msr_normalop* op) {
    const int pad = op->cart ? 1 : 2;
    GridDims g{{pad * op->N[0], pad * op->N[1], op->d == 3 ? pad * op->N[2] : 1}, op->cart};
    return g;
}

int64_t op_grid_cells(const msr_normalop* op) {
    GridDims g = op_grid(op);
    return g.n[0] * g.n[1] * g.n[2];
}

// Fill the dense kernel from a packed (P, Nkm) kernel and 1-based kmask.
void op_load_packed(msr_normalop* op, const void* lambda, const int64_t* kmask, int64_t nkm,
                    hipStream_t st) {
    const int64_t G = op_grid_cells(op);
    const size_t es = op->lam_real ? sizeof(float) : sizeof(float2);
    op->lam->alloc(es * (size_t)op->P * G);
    fill_floats(op->lam->p, (int64_t)(op->lam->bytes / sizeof(float)), 0.f, st);
    op->nkm = nkm;
    op->kmask->alloc(sizeof(int64_t) * std::max<int64_t>(nkm, 1));
    if (nkm > 0)
        MSR_HIP(hipMemcpyAsync(op->kmask->p, kmask, sizeof(int64_t) * nkm, hipMemcpyDeviceToDevice, st));
    DevBuf bad;
    bad.alloc(sizeof(int));
    fill_floats(bad.p, 1, 0.f, st);
    if (nkm > 0) {
        int grid = (int)std::min<int64_t>((nkm * op->P + 255) / 256, 16384);
        GridDims g = op_grid(op);
        if (op->lam_real)
            hipLaunchKernelGGL(scatter_kernel<float>, dim3(grid), dim3(256), 0, st,
                               (const float*)lambda, op->kmask->as<int64_t>(), nkm, op->P, op->d, g,
                               op->lam->as<float>(), bad.as<int>());
        else
            hipLaunchKernelGGL(scatter_kernel<float2>, dim3(grid), dim3(256), 0, st,
                               (const float2*)lambda, op->kmask->as<int64_t>(), nkm, op->P, op->d, g,
                               op->lam->as<float2>(), bad.as<int>());
        MSR_CHECK_LAUNCH();
    }
    int hbad = 0;
    MSR_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    MSR_REQUIRE(hbad == 0, MSR_ERR_ASSERTION,
                "kmask_indcs out of bounds: entries must lie in 1..prod(2 .* img_shape) = %lld",
                (long long)G);
}

msr_normalop* op_new(msr_ctx* ctx, int d, const int64_t* N, int K, int lam_real, int cart) {
    if (cart) check_shape_cart(d, N);
    else check_shape(d, N);
    MSR_REQUIRE(K >= 1, MSR_ERR_ARGUMENT, "K = %d coefficients", K);
    MSR_REQUIRE(K <= (cart ? 8 : MAX_COEFFS), MSR_ERR_UNSUPPORTED,
                "K = %d coefficients, supported 1..%d", K, cart ? 8 : MAX_COEFFS);
    auto op = std::make_unique<msr_normalop>();
    op->ctx = ctx;
    op->d = d;
    op->V = 1;
    for (int i = 0; i < 3; ++i) op->N[i] = i < d ? N[i] : 1;
    for (int i = 0; i < d; ++i) op->V *= N[i];
    op->K = K;
    op->P = cart ? K * K : K * (K + 1) / 2;
    op->lam_real = lam_real;
    op->cart = cart ? 1 : 0;
    // MSR_APPLY_LINES=1 sends every non-Cartesian operator to the line-kernel
    // apply (a cross-check of the two paths on shapes both cover)
    op->fused = !cart && fused_shape(d, N, K) && env_int("MSR_APPLY_LINES", 0) == 0;
    return op.release();
}

// ------------------------------------------------------------------ applies
Apply2DArgs args2d(msr_normalop* op, int am) {
    const bool f64 = am != AM_F32;
    Apply2DArgs a;
    a.nx = (int)op->N[0];
    a.ny = (int)op->N[1];
    a.K = op->K;
    a.C = op->C;
    a.f64 = f64 ? 1 : 0;
    a.compact = am == AM_F64C ? 1 : 0;
    a.cmaps = op->identity_cmaps ? nullptr : op->cmaps.as<float2>();
    a.lam = op->lam->p;
    a.lam_real = op->lam_real;
    a.T = op->T.p;
    a.partial = op->partial.p;
    a.G = op->G;
    a.Gb = op->knobs[am].Gb;
    a.order_a = op->knobs[am].order_a;
    a.order_c = op->knobs[am].order_c;
    if (f64) {
        a.tw_x = twiddle_table64(op->ctx, 2 * a.nx);
        a.tw_y = twiddle_table64(op->ctx, 2 * a.ny);
    } else {
        a.tw_x = twiddle_table(op->ctx, 2 * a.nx);
        a.tw_y = twiddle_table(op->ctx, 2 * a.ny);
    }
    a.b = nullptr;
    a.skip = nullptr;
    return a;
}

double op_scale(const msr_normalop* op) { return 1.0 / (double)op_grid_cells(op); }

// partial[g] <- unscaled group sums of A^H A b (b and partial of the apply type)
// am: ApplyMode.  The compact storage of AM_F64C applies to the fused 2D and
// 3D passes; the Cartesian and line-kernel passes keep ComplexF64
// intermediates in that mode.
void op_passes(msr_normalop* op, const void* b, int am, hipStream_t st, const int* skip) {
    const bool f64 = am != AM_F32;
    MSR_REQUIRE(op->scratch_es >= (f64 ? sizeof(double2) : sizeof(float2)), MSR_ERR_ERROR,
                "apply scratch not allocated for this precision");
    if (op->cart) {
        launch_cart_passes(op, b, f64, st, skip);
    } else if (!op->fused) {
        launch_lines_passes(op, b, f64, st, skip);
    } else if (op->d == 2) {
        Apply2DArgs a = args2d(op, am);
        a.b = b;
        a.skip = skip;
        launch_apply2d_passes(a, st);
    } else {
        launch_apply3d_passes(op, b, am, st, skip);
    }
}

void op_apply(msr_normalop* op, float2* x, const float2* b, float2 alpha, float2 beta,
              bool beta_zero, hipStream_t st) {
    op_passes(op, b, AM_F32, st, nullptr);
    launch_epilogue<float2>(op->partial.as<float2>(), op->G, op->V * op->K, x, alpha, beta,
                            beta_zero, op_scale(op), st, nullptr);
}

void op_apply_f64(msr_normalop* op, double2* x, const double2* b, double2 alpha, double2 beta,
                  bool beta_zero, hipStream_t st, int am) {
    op_ensure_scratch(op, true);
    op_passes(op, b, am, st, nullptr);
    launch_epilogue<double2>(op->partial.as<double2>(), op->G, op->V * op->K, x, alpha, beta,
                             beta_zero, op_scale(op), st, nullptr);
}

// The reference GPU order: one coil at a time, kernel re-read per coil.
void op_apply_percoil(msr_normalop* op, float2* x, const float2* b, float2 alpha, float2 beta,
                      bool beta_zero, hipStream_t st) {
    MSR_REQUIRE(!op->cart, MSR_ERR_UNSUPPORTED,
                "the per-coil apply order exists for the non-Cartesian operator only");
    launch_scale<float2>(x, op->V * op->K, beta, beta_zero, st);
    for (int c = 0; c < op->C; ++c) {
        if (!op->fused) {
            launch_lines_passes(op, b, false, st, nullptr, c, c + 1);
        } else if (op->d == 2) {
            Apply2DArgs a = args2d(op, AM_F32);
            a.C = 1;
            a.G = 1;
            a.Gb = 1;
            if (a.cmaps) a.cmaps += (size_t)c * op->V;
            a.b = b;
            launch_apply2d_passes(a, st);
        } else {
            launch_apply3d_passes_coil(op, b, c, st);
        }
        launch_epilogue<float2>(op->partial.as<float2>(), 1, op->V * op->K, x, alpha,
                                make_float2(1.f, 0.f), false, op_scale(op), st, nullptr);
    }
}

void op_export(const msr_normalop* op, void* lambda_out, int64_t* kmask_out, hipStream_t st) {
    if (op->nkm == 0) return;
    if (kmask_out)
        MSR_HIP(hipMemcpyAsync(kmask_out, op->kmask->p, sizeof(int64_t) * op->nkm,
                               hipMemcpyDeviceToDevice, st));
    if (lambda_out) {
        int grid = (int)std::min<int64_t>((op->nkm * op->P + 255) / 256, 16384);
        GridDims g = op_grid(op);
        if (op->lam_real)
            hipLaunchKernelGGL(gather_kernel<float>, dim3(grid), dim3(256), 0, st,
                               op->lam->as<float>(), op->kmask->as<int64_t>(), op->nkm, op->P, op->d,
                               g, (float*)lambda_out);
        else
            hipLaunchKernelGGL(gather_kernel<float2>, dim3(grid), dim3(256), 0, st,
                               op->lam->as<float2>(), op->kmask->as<int64_t>(), op->nkm, op->P,
                               op->d, g, (float2*)lambda_out);
        MSR_CHECK_LAUNCH();
    }
}

// ------------------------------------------------------------------ CG
// Buffers of a CG in precision prec (MSR_PREC_*): r, u (and the FP64 x) in the
// vector type, c in the apply type, and an FP32 copy of u when the vectors are
// FP64 but the apply is FP32.  Also sizes the apply scratch.
size_t op_footprint(const msr_normalop* op, int prec) {
    const bool v64 = prec != MSR_PREC_F32, a64 = apply_mode(prec) != AM_F32;
    const size_t n = (size_t)op->V * op->K;
    const size_t es = a64 ? sizeof(double2) : sizeof(float2);  // apply intermediates
    const size_t vs = v64 ? sizeof(double2) : sizeof(float2);  // CG vectors
    const int64_t nx = op->N[0], ny = op->N[1], nz = op->N[2];
    size_t scratch = es * (size_t)op->G * op->K * op->V;
    if (op->cart) scratch += 2 * es * (size_t)op->C * op->K * op->V;
    else if (!op->fused) scratch += 2 * lines_scratch_bytes(op, es);
    else if (op->d == 2) scratch += es * (size_t)op->C * op->K * (2 * nx) * ny;
    else scratch += es * (size_t)op->C * op->K * nz * ny * 2 * nx * 3;  // T1 + T2
    size_t cg = vs * n * 2 + es * n;                                   // r, u, c
    if (v64) cg += sizeof(double2) * n;                                // x in FP64
    if (prec == MSR_PREC_MIXED) cg += sizeof(float2) * n;
    cg += sizeof(CgScalars) + sizeof(double) * CG_PART_DOUBLES;
    const size_t maps = op->identity_cmaps ? 0 : sizeof(float2) * (size_t)op->C * op->V;
    return maps + std::max(scratch, (size_t)op->T.bytes + op->T2.bytes + op->partial.bytes) + cg;
}

size_t op_kernel_bytes(const msr_normalop* op) {
    return (op->lam ? op->lam->bytes : 0) + (op->kmask ? op->kmask->bytes : 0);
}

void op_cg_buffers(msr_normalop* op, int prec) {
    const size_t n = (size_t)op->V * op->K;
    const bool v64 = prec != MSR_PREC_F32, a64 = apply_mode(prec) != AM_F32;
    const size_t es = v64 ? sizeof(double2) : sizeof(float2);
    op->cg_r.alloc(es * n);
    op->cg_u.alloc(es * n);
    op->cg_c.alloc((a64 ? sizeof(double2) : sizeof(float2)) * n);
    if (v64) op->cg_x64.alloc(sizeof(double2) * n);
    if (prec == MSR_PREC_MIXED) op->cg_u32.alloc(sizeof(float2) * n);
    op->cg_scal.alloc(sizeof(CgScalars));
    op->cg_part.alloc(sizeof(double) * CG_PART_DOUBLES);
    op_ensure_scratch(op, a64);
}

template <class VT, class AT>
CgVecs<VT, AT> op_cg_vecs(msr_normalop* op, float2* x) {
    CgVecs<VT, AT> v;
    v.r = op->cg_r.as<VT>();
    v.u = op->cg_u.as<VT>();
    v.c = op->cg_c.as<AT>();
    if constexpr (std::is_same<VT, double2>::value) v.x = op->cg_x64.as<double2>();
    else v.x = x;
    if constexpr (std::is_same<VT, AT>::value) v.ua = v.u;
    else v.ua = op->cg_u32.as<AT>();
    return v;
}

template CgVecs<float2, float2> op_cg_vecs<float2, float2>(msr_normalop*, float2*);
template CgVecs<double2, float2> op_cg_vecs<double2, float2>(msr_normalop*, float2*);
template CgVecs<double2, double2> op_cg_vecs<double2, double2>(msr_normalop*, float2*);

// One key per captured CG graph: everything the captured kernels hold.
std::vector<const void*> op_graph_key(const msr_normalop* op) {
    std::vector<const void*> key;
    for (const DevBuf* d : {&op->cg_r, &op->cg_u, &op->cg_c, &op->cg_scal, &op->cg_part,
                            &op->cg_x64, &op->cg_u32, &op->T, &op->T2, &op->partial})
        key.push_back(d->p);
    return key;
}

__global__ void f32_to_f64_kernel(const float2* __restrict__ a, double2* __restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = make_double2(a[i].x, a[i].y);
}

// one CG iteration with device scalars S and reduction partials part (all
// kernels skip once done is set)
template <class VT, class AT>
static void cg_iteration(msr_normalop* op, const CgVecs<VT, AT>& v, CgScalars* S, double* part,
                         int am, hipStream_t st) {
    const int64_t n = op->V * op->K;
    launch_cg_u(v, n, S, st);
    op_passes(op, v.ua, am, st, &S->done);
    launch_cg_reduce_dot(op->partial.as<AT>(), op->G, n, v, op_scale(op), part, S, st, false);
    launch_cg_xr_step(v, n, part, S, st);
}

// c = A x0 in the apply precision
template <class VT, class AT>
static void cg_apply_x0(msr_normalop* op, const CgVecs<VT, AT>& v, const float2* x, int am,
                        hipStream_t st) {
    if constexpr (std::is_same<AT, double2>::value) {
        const int64_t n = op->V * op->K;
        int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(f32_to_f64_kernel, dim3(grid), dim3(256), 0, st, x, v.ua, n);
        MSR_CHECK_LAUNCH();
        op_passes(op, v.ua, am, st, nullptr);
        launch_epilogue<double2>(op->partial.as<double2>(), op->G, n, v.c, make_double2(1.0, 0.0),
                                 make_double2(0.0, 0.0), true, op_scale(op), st, nullptr);
    } else {
        op_apply(op, v.c, x, make_float2(1.f, 0.f), make_float2(0.f, 0.f), true, st);
    }
}

template <class VT, class AT>
static void cg_prologue(msr_normalop* op, const CgVecs<VT, AT>& v, CgScalars* S, double* part,
                        float2* x, const float2* b, int maxiter, double reltol, double abstol,
                        bool initially_zero, int am, hipStream_t st) {
    const int64_t n = op->V * op->K;
    if (initially_zero) {
        launch_cg_start(b, x, v, n, true, part, S, reltol, abstol, maxiter, false, st);
    } else {
        cg_apply_x0(op, v, x, am, st);
        launch_cg_start(b, x, v, n, false, part, S, reltol, abstol, maxiter, true, st);
    }
}

template <class VT, class AT>
static void cg_epilogue(msr_normalop* op, const CgVecs<VT, AT>& v, float2* x, hipStream_t st) {
    if constexpr (std::is_same<VT, double2>::value) launch_to_f32(v.x, x, op->V * op->K, st);
}

template <class VT, class AT>
static void cg_run(msr_normalop* op, float2* x, const float2* b, int maxiter, double reltol,
                   double abstol, bool initially_zero, int mode, int prec, hipStream_t st) {
    const CgVecs<VT, AT> v = op_cg_vecs<VT, AT>(op, x);
    auto* S = op->cg_scal.as<CgScalars>();
    double* part = op->cg_part.as<double>();
    const int am = apply_mode(prec);
    if (mode == MSR_CG_DEVICE) {
        cg_prologue(op, v, S, part, x, b, maxiter, reltol, abstol, initially_zero, am, st);
        for (int it = 0; it < maxiter; ++it) cg_iteration(op, v, S, part, am, st);
        cg_epilogue(op, v, x, st);
    } else if (mode == MSR_CG_GRAPH) {
        // the captured kernels hold the buffer addresses: they are part of the
        // key (another call may re-allocate them for another precision)
        std::vector<const void*> key = op_graph_key(op);
        key.insert(key.end(), {x, b, (const void*)(intptr_t)(initially_zero ? 1 : 2), st});
        const bool hit = op->graph && op->graph_key == key && op->graph_maxiter == maxiter &&
                         op->graph_tol[0] == reltol && op->graph_tol[1] == abstol &&
                         op->graph_fp64 == prec;
        if (!hit) {
            if (op->graph) {
                MSR_HIP(hipStreamSynchronize(op->graph_stream));  // a replay may still run
                (void)hipGraphExecDestroy(op->graph);
                op->graph = nullptr;
            }
            MSR_REQUIRE(st != nullptr, MSR_ERR_ARGUMENT,
                        "MSR_CG_GRAPH needs an explicit (non-default) stream");
            hipGraph_t g = nullptr;
            MSR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                cg_prologue(op, v, S, part, x, b, maxiter, reltol, abstol, initially_zero, am, st);
                for (int it = 0; it < maxiter; ++it) cg_iteration(op, v, S, part, am, st);
                cg_epilogue(op, v, x, st);
            } catch (...) {
                (void)hipStreamEndCapture(st, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            MSR_HIP(hipStreamEndCapture(st, &g));
            hipGraphExec_t ge = nullptr;
            hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            MSR_HIP(e);
            op->graph = ge;
            op->graph_key = key;
            op->graph_stream = st;
            op->graph_maxiter = maxiter;
            op->graph_tol[0] = reltol;
            op->graph_tol[1] = abstol;
            op->graph_fp64 = prec;
        }
        MSR_HIP(hipGraphLaunch(op->graph, st));
    } else if (mode == MSR_CG_HOST) {
        // IterativeSolvers-style: the step length and the residual norm are
        // read back to the host every iteration; applies run coil by coil
        // (the Cartesian operator, which the reference has on the CPU only,
        // applies all coils at once).
        if constexpr (std::is_same<AT, float2>::value) {
            const int64_t n = op->V * op->K;
            cg_prologue(op, v, S, part, x, b, maxiter, reltol, abstol, initially_zero, am, st);
            CgScalars hs;
            for (int it = 0; it < maxiter; ++it) {
                MSR_HIP(hipMemcpyAsync(&hs, S, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
                MSR_HIP(hipStreamSynchronize(st));
                if (hs.done) break;
                launch_cg_u(v, n, S, st);
                if (op->cart)
                    op_apply(op, v.c, v.ua, make_float2(1.f, 0.f), make_float2(0.f, 0.f), true, st);
                else
                    op_apply_percoil(op, v.c, v.ua, make_float2(1.f, 0.f), make_float2(0.f, 0.f),
                                     true, st);
                launch_cg_reduce_dot(v.c, 1, n, v, 1.0, part, S, st);
                MSR_HIP(hipMemcpyAsync(&hs, S, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
                MSR_HIP(hipStreamSynchronize(st));
                launch_cg_xr(v, n, part, S, st);
            }
            cg_epilogue(op, v, x, st);
        } else {
            fail(MSR_ERR_UNSUPPORTED, "MSR_CG_HOST runs the reference-order FP32 apply; "
                                      "use MSR_PREC_F32 or MSR_PREC_MIXED");
        }
    } else {
        fail(MSR_ERR_ARGUMENT, "unknown CG mode %d", mode);
    }
    (void)prec;
}

void op_cg(msr_normalop* op, float2* x, const float2* b, int maxiter, double reltol,
           double abstol, bool initially_zero, int mode, int prec, int* iters_out,
           double* res_out, hipStream_t st) {
    MSR_REQUIRE(prec_valid(prec), MSR_ERR_ARGUMENT, "unknown CG precision %d", prec);
    op_cg_buffers(op, prec);
    if (prec == MSR_PREC_F64 || prec == MSR_PREC_F64C)
        cg_run<double2, double2>(op, x, b, maxiter, reltol, abstol, initially_zero, mode, prec, st);
    else if (prec == MSR_PREC_MIXED)
        cg_run<double2, float2>(op, x, b, maxiter, reltol, abstol, initially_zero, mode, prec, st);
    else
        cg_run<float2, float2>(op, x, b, maxiter, reltol, abstol, initially_zero, mode, prec, st);
    if (iters_out || res_out) {
        CgScalars hs;
        MSR_HIP(hipMemcpyAsync(&hs, op->cg_scal.p, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
        MSR_HIP(hipStreamSynchronize(st));
        if (iters_out) *iters_out = hs.iter;
        if (res_out) *res_out = hs.res;
    }
}

// ------------------------------------------------- CG without allocation
// Device memory of msr_cg_noalloc: the CG vectors, scalars and reduction
// partials of one solve, carved from one array of CG_NOALLOC_BYTES that each
// device holds from the library's load to the process end.  In host code the
// array's name stands for a host-side shadow, not a device address: the
// device address is looked up once per device (cg_noalloc_setup, from
// msr_ctx_create with the device current) and only that address is used.

}  // namespace msr

static __device__ __align__(256) unsigned char scratch[msr::CG_NOALLOC_BYTES];

namespace msr {

static std::mutex g_noalloc_mu;
static std::map<int, unsigned char*> g_noalloc_ptr;  // device -> address of `scratch`

void cg_noalloc_setup(int device) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr.count(device)) return;
    void* p = nullptr;
    MSR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MSR_REQUIRE(p != nullptr, MSR_ERR_HIP, "no device address for the CG scratch array");
    g_noalloc_ptr[device] = static_cast<unsigned char*>(p);
}

static unsigned char* noalloc_base(int device) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(device);
    MSR_REQUIRE(it != g_noalloc_ptr.end(), MSR_ERR_ERROR,
                "CG scratch of device %d not set up (msr_ctx_create does it)", device);
    return it->second;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch a solve of n unknowns at precision prec takes.
size_t cg_noalloc_bytes(int64_t n, int prec) {
    const bool v64 = prec != MSR_PREC_F32, a64 = apply_mode(prec) != AM_F32;
    const size_t vs = v64 ? sizeof(double2) : sizeof(float2);
    const size_t es = a64 ? sizeof(double2) : sizeof(float2);
    size_t b = 2 * align256(vs * n) + align256(es * n);        // r, u, c
    if (v64) b += align256(sizeof(double2) * n);               // x in FP64
    if (prec == MSR_PREC_MIXED) b += align256(sizeof(float2) * n);
    return b + align256(sizeof(CgScalars)) + align256(sizeof(double) * CG_PART_DOUBLES);
}

template <class VT, class AT>
static void cg_noalloc_run(msr_normalop* op, unsigned char* base, float2* x, const float2* b,
                           int maxiter, double reltol, double abstol, bool initially_zero, int am,
                           hipStream_t st, CgScalars** S_out) {
    const size_t n = (size_t)op->V * op->K;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += align256(bytes);
        return p;
    };
    CgVecs<VT, AT> v;
    v.r = reinterpret_cast<VT*>(take(sizeof(VT) * n));
    v.u = reinterpret_cast<VT*>(take(sizeof(VT) * n));
    v.c = reinterpret_cast<AT*>(take(sizeof(AT) * n));
    if constexpr (std::is_same<VT, double2>::value) v.x = reinterpret_cast<double2*>(take(sizeof(double2) * n));
    else v.x = x;
    if constexpr (std::is_same<VT, AT>::value) v.ua = v.u;
    else v.ua = reinterpret_cast<AT*>(take(sizeof(AT) * n));
    auto* S = reinterpret_cast<CgScalars*>(take(sizeof(CgScalars)));
    auto* part = reinterpret_cast<double*>(take(sizeof(double) * CG_PART_DOUBLES));
    // every buffer above is written by the start kernels before it is read
    // (r, u, x and the scalars by cg_start; c, part by each iteration), so
    // nothing carried over from an earlier call is used
    cg_prologue(op, v, S, part, x, b, maxiter, reltol, abstol, initially_zero, am, st);
    for (int it = 0; it < maxiter; ++it) cg_iteration(op, v, S, part, am, st);
    cg_epilogue(op, v, x, st);
    *S_out = S;
}

void op_cg_noalloc(msr_normalop* op, float2* x, const float2* b, int maxiter, double reltol,
                   double abstol, bool initially_zero, int prec, int* iters_out, double* res_out,
                   hipStream_t st) {
    MSR_REQUIRE(prec_valid(prec), MSR_ERR_ARGUMENT, "unknown CG precision %d", prec);
    MSR_REQUIRE(maxiter >= 0, MSR_ERR_ARGUMENT, "maxiter = %d", maxiter);
    const size_t need = cg_noalloc_bytes(op->V * op->K, prec);
    MSR_REQUIRE(need <= CG_NOALLOC_BYTES, MSR_ERR_ARGUMENT,
                "the solve needs %zu bytes of CG state, above the %zu bytes this entry point "
                "holds (use msr_cg)", need, (size_t)CG_NOALLOC_BYTES);
    const int am = apply_mode(prec);
    const size_t es = am != AM_F32 ? sizeof(double2) : sizeof(float2);
    MSR_REQUIRE(op->scratch_es >= es, MSR_ERR_ARGUMENT,
                "the operator's apply scratch is not sized for this precision yet (this entry "
                "point allocates nothing: run one msr_normalop_apply_f64 or msr_cg at this "
                "precision first)");
    unsigned char* base = noalloc_base(op->ctx->device);
    CgScalars* S = nullptr;
    if (am != AM_F32)
        cg_noalloc_run<double2, double2>(op, base, x, b, maxiter, reltol, abstol, initially_zero,
                                         am, st, &S);
    else if (prec == MSR_PREC_MIXED)
        cg_noalloc_run<double2, float2>(op, base, x, b, maxiter, reltol, abstol, initially_zero,
                                        am, st, &S);
    else
        cg_noalloc_run<float2, float2>(op, base, x, b, maxiter, reltol, abstol, initially_zero,
                                       am, st, &S);
    if (iters_out || res_out) {
        CgScalars hs;
        MSR_HIP(hipMemcpyAsync(&hs, S, sizeof(CgScalars), hipMemcpyDeviceToHost, st));
        MSR_HIP(hipStreamSynchronize(st));
        if (iters_out) *iters_out = hs.iter;
        if (res_out) *res_out = hs.res;
    }
}

}  // namespace msr

msr_normalop::~msr_normalop() {
    if (graph) (void)hipGraphExecDestroy(graph);
}
