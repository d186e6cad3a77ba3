// General Toeplitz apply on the line kernels: the normal operator of
// ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:193-218 for every grid the line
// FFTs cover (each 2N 2,3,5-smooth up to 2048) and any number of coefficients
// K, in 2D and 3D.  The fused passes (apply2d*.hip, apply3d*.hip) cover the
// common shapes (2N in the fused length set, up to 512, K <= 8) faster; this
// path takes the rest.
//
// For a chunk of Cb coils at a time, with the grids of the chunk in two
// ping-pong buffers A and B of Cb K prod(2N) elements:
//   expand   A[c, k] = S_c . b_k                          (image grid)
//   x rows   zero-padded FFT, Nx inputs -> 2Nx bins
//   y (z) columns, the same                                 -> A or B: [c, k][2N grid]
//   multiply out[c, j](q) = sum_k Lambda_jk(q) in[c, k](q)  (packed Hermitian or
//            real symmetric, Lambda_kj = conj(Lambda_jk); zero off the kmask)
//   z (3D), y columns: inverse FFT, first N bins kept (crop)
//   x rows: inverse FFT, first Nx bins kept
//   combine  partial[k] (+)= sum_c conj(S_c) . img[c, k]
// The FFTs are unnormalised; the epilogue applies 1 / prod(2N) as for the
// fused passes.  The arithmetic and the grids are ComplexF32 or ComplexF64
// (the compact 8-byte storage of the fused 2D passes maps to ComplexF64 here).
#include <algorithm>

#include "fftlines.hpp"
#include "normalop.hpp"

namespace msr {

// A[(cc K + k) V + v] = S_{c0+cc}[v] b[k V + v]
template <class AT>
__global__ __launch_bounds__(256) void lines_expand_kernel(const AT* __restrict__ b,
                                                           const float2* __restrict__ cmaps,
                                                           int c0, int nc, int K, int64_t V,
                                                           AT* __restrict__ A,
                                                           const int* __restrict__ skip) {
    if (skip && *skip) return;
    using RT = real_t<AT>;
    const int64_t n = (int64_t)nc * K * V;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i % V;
        const int64_t ck = i / V;
        const int k = (int)(ck % K);
        const int c = c0 + (int)(ck / K);
        AT x = b[(int64_t)k * V + v];
        if (cmaps) {
            const float2 s = cmaps[(int64_t)c * V + v];
            x = cmul(AT{(RT)s.x, (RT)s.y}, x);
        }
        A[i] = x;
    }
}

// Offset of packed entry p of grid cell (qx, qy, qz) in the dense kernel:
// 2D lam[(qx P + p) 2Ny + qy], 3D lam[((qy 2Nx + qx) P + p) 2Nz + qz].
__device__ __forceinline__ int64_t lines_lam_offset(int d, int64_t nx2, int64_t ny2, int64_t nz2,
                                                    int P, int64_t qx, int64_t qy, int64_t qz,
                                                    int p) {
    if (d == 2) return (qx * P + p) * ny2 + qy;
    return ((qy * nx2 + qx) * P + p) * nz2 + qz;
}

// out[(cc K + j) G + q] = sum_k Lambda_jk(q) in[(cc K + k) G + q], one thread
// per (coil of the chunk, cell, j); Lambda from the packed upper triangle
// (entry j + k (k + 1) / 2 for j <= k), promoted to the arithmetic type.
template <class AT, class LT>
__global__ __launch_bounds__(256) void lines_multiply_kernel(
    const AT* __restrict__ in, AT* __restrict__ out, const LT* __restrict__ lam, int d,
    int64_t nx2, int64_t ny2, int64_t nz2, int K, int nc, const int* __restrict__ skip) {
    if (skip && *skip) return;
    using RT = real_t<AT>;
    const int P = K * (K + 1) / 2;
    const int64_t G = nx2 * ny2 * nz2;
    const int64_t n = (int64_t)nc * K * G;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = i % G;
        const int64_t cj = i / G;
        const int j = (int)(cj % K);
        const int64_t cc = cj / K;
        const int64_t qx = q % nx2, r = q / nx2;
        const int64_t qy = r % ny2, qz = r / ny2;
        const AT* x = in + cc * K * G + q;
        AT acc{0, 0};
        for (int k = 0; k < K; ++k) {
            const bool upper = j <= k;
            const int p = upper ? j + k * (k + 1) / 2 : k + j * (j + 1) / 2;
            const LT l = lam[lines_lam_offset(d, nx2, ny2, nz2, P, qx, qy, qz, p)];
            const AT xk = x[(int64_t)k * G];
            if constexpr (sizeof(LT) == sizeof(float)) {
                acc.x = fma((RT)l, xk.x, acc.x);
                acc.y = fma((RT)l, xk.y, acc.y);
            } else {
                const AT la{(RT)l.x, (RT)l.y};
                if (upper) cfma(acc, la, xk);
                else cfmac(acc, la, xk);  // Lambda_jk = conj(Lambda_kj)
            }
        }
        out[i] = acc;
    }
}

// partial[k V + v] (+)= sum_cc conj(S_{c0+cc}[v]) img[(cc K + k) V + v]
template <class AT>
__global__ __launch_bounds__(256) void lines_combine_kernel(const AT* __restrict__ img,
                                                            const float2* __restrict__ cmaps,
                                                            int c0, int nc, int K, int64_t V,
                                                            int accumulate,
                                                            AT* __restrict__ partial,
                                                            const int* __restrict__ skip) {
    if (skip && *skip) return;
    using RT = real_t<AT>;
    const int64_t n = (int64_t)K * V;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i % V;
        const int k = (int)(i / V);
        AT acc = accumulate ? partial[i] : AT{0, 0};
        for (int cc = 0; cc < nc; ++cc) {
            const AT x = img[((int64_t)cc * K + k) * V + v];
            if (cmaps) {
                const float2 s = cmaps[(int64_t)(c0 + cc) * V + v];
                cfmac(acc, AT{(RT)s.x, (RT)s.y}, x);
            } else {
                acc = cadd(acc, x);
            }
        }
        partial[i] = acc;
    }
}

static unsigned lines_grid(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384));
}

int lines_chunk_coils(const msr_normalop* op) {
    // the two grids of a chunk in ComplexF64 stay within 2 x 1 GiB (at least
    // one coil per chunk)
    int64_t g = op->K;
    for (int a = 0; a < op->d; ++a) g *= 2 * op->N[a];
    const int64_t per = g * (int64_t)sizeof(double2);
    const int64_t cb = std::max<int64_t>(1, ((int64_t)1 << 30) / per);
    return (int)std::min<int64_t>(cb, op->C);
}

size_t lines_scratch_bytes(const msr_normalop* op, size_t es) {
    int64_t g = (int64_t)op->Cb * op->K;
    for (int a = 0; a < op->d; ++a) g *= 2 * op->N[a];
    return es * (size_t)g;
}

template <class AT>
static void lines_passes_t(msr_normalop* op, const AT* b, int c_begin, int c_end, hipStream_t st,
                           const int* skip) {
    const int d = op->d, K = op->K;
    const int64_t nx = op->N[0], ny = op->N[1], nz = d == 3 ? op->N[2] : 1;
    const int64_t nx2 = 2 * nx, ny2 = 2 * ny, nz2 = d == 3 ? 2 * nz : 1;
    const int64_t V = op->V, G = nx2 * ny2 * nz2;
    const float2* cm = op->identity_cmaps ? nullptr : op->cmaps.as<float2>();
    const bool f64 = sizeof(AT) == sizeof(double2);
    AT* bufA = op->T.as<AT>();
    AT* bufB = op->T2.as<AT>();
    AT* partial = op->partial.as<AT>();
    MSR_REQUIRE(op->T.bytes >= lines_scratch_bytes(op, sizeof(AT)) &&
                    op->T2.bytes >= lines_scratch_bytes(op, sizeof(AT)),
                MSR_ERR_ERROR, "apply scratch not allocated for this precision");
    auto line = [&](int nf, int dir, int lin, int lout) {
        LinePass p;
        p.f64 = f64 ? 1 : 0;
        p.nf = nf;
        p.dir = dir;
        p.lin = lin;
        p.lout = lout;
        p.sel = SEL_FIRST;
        p.isel = SEL_FIRST;
        p.skip = skip;
        return p;
    };
    for (int c0 = c_begin; c0 < c_end; c0 += op->Cb) {
        const int nc = std::min(op->Cb, c_end - c0);
        const int64_t nck = (int64_t)nc * K;
        hipLaunchKernelGGL(lines_expand_kernel<AT>, dim3(lines_grid(nck * V)), dim3(256), 0, st, b,
                           cm, c0, nc, K, V, bufA, skip);
        MSR_CHECK_LAUNCH();
        // x rows: A [ck][z][y][x] -> B [ck][z][y][qx]
        LinePass px = line((int)nx2, -1, (int)nx, (int)nx2);
        px.in = bufA;
        px.out = bufB;
        px.nlines = nck * nz * ny;
        px.in_pitch = nx;
        px.out_pitch = nx2;
        run_rows(px, op->ctx, st);
        // y columns: B [ck][z][y][qx] -> A [ck][z][qy][qx]
        LinePass py = line((int)ny2, -1, (int)ny, (int)ny2);
        py.in = bufB;
        py.out = bufA;
        py.nb = nck * nz;
        py.ncols = nx2;
        py.in_bstride = ny * nx2;
        py.in_stride = nx2;
        py.out_bstride = ny2 * nx2;
        py.out_stride = nx2;
        run_cols(py, op->ctx, st);
        AT* cur = bufA;
        AT* other = bufB;
        if (d == 3) {
            // z columns: A [ck][z][qy][qx] -> B [ck][qz][qy][qx]
            LinePass pz = line((int)nz2, -1, (int)nz, (int)nz2);
            pz.in = bufA;
            pz.out = bufB;
            pz.nb = nck;
            pz.ncols = ny2 * nx2;
            pz.in_bstride = nz * ny2 * nx2;
            pz.in_stride = ny2 * nx2;
            pz.out_bstride = G;
            pz.out_stride = ny2 * nx2;
            run_cols(pz, op->ctx, st);
            std::swap(cur, other);
        }
        // per-cell K x K multiply: cur -> other
        if (op->lam_real)
            hipLaunchKernelGGL((lines_multiply_kernel<AT, float>), dim3(lines_grid(nck * G)),
                               dim3(256), 0, st, cur, other, op->lam->as<float>(), d, nx2, ny2,
                               nz2, K, nc, skip);
        else
            hipLaunchKernelGGL((lines_multiply_kernel<AT, float2>), dim3(lines_grid(nck * G)),
                               dim3(256), 0, st, cur, other, op->lam->as<float2>(), d, nx2, ny2,
                               nz2, K, nc, skip);
        MSR_CHECK_LAUNCH();
        std::swap(cur, other);
        if (d == 3) {
            // z columns, inverse, crop: [ck][qz][qy][qx] -> [ck][z][qy][qx]
            LinePass iz = line((int)nz2, +1, (int)nz2, (int)nz);
            iz.in = cur;
            iz.out = other;
            iz.nb = nck;
            iz.ncols = ny2 * nx2;
            iz.in_bstride = G;
            iz.in_stride = ny2 * nx2;
            iz.out_bstride = nz * ny2 * nx2;
            iz.out_stride = ny2 * nx2;
            run_cols(iz, op->ctx, st);
            std::swap(cur, other);
        }
        // y columns, inverse, crop: [ck][z][qy][qx] -> [ck][z][y][qx]
        LinePass iy = line((int)ny2, +1, (int)ny2, (int)ny);
        iy.in = cur;
        iy.out = other;
        iy.nb = nck * nz;
        iy.ncols = nx2;
        iy.in_bstride = ny2 * nx2;
        iy.in_stride = nx2;
        iy.out_bstride = ny * nx2;
        iy.out_stride = nx2;
        run_cols(iy, op->ctx, st);
        std::swap(cur, other);
        // x rows, inverse, crop: [ck][z][y][qx] -> [ck][z][y][x]
        LinePass ix = line((int)nx2, +1, (int)nx2, (int)nx);
        ix.in = cur;
        ix.out = other;
        ix.nlines = nck * nz * ny;
        ix.in_pitch = nx2;
        ix.out_pitch = nx;
        run_rows(ix, op->ctx, st);
        std::swap(cur, other);
        hipLaunchKernelGGL(lines_combine_kernel<AT>, dim3(lines_grid((int64_t)K * V)), dim3(256),
                           0, st, cur, cm, c0, nc, K, V, c0 > c_begin ? 1 : 0, partial, skip);
        MSR_CHECK_LAUNCH();
    }
}

void launch_lines_passes(msr_normalop* op, const void* b, bool f64, hipStream_t st,
                         const int* skip, int c_begin, int c_end) {
    if (c_end < 0) c_end = op->C;
    MSR_REQUIRE(0 <= c_begin && c_begin < c_end && c_end <= op->C, MSR_ERR_ERROR,
                "coil range [%d, %d) of %d coils", c_begin, c_end, op->C);
    if (f64) lines_passes_t<double2>(op, static_cast<const double2*>(b), c_begin, c_end, st, skip);
    else lines_passes_t<float2>(op, static_cast<const float2*>(b), c_begin, c_end, st, skip);
}

}  // namespace msr
