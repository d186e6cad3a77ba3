// Fused 3D Toeplitz subspace normal operator: five streaming passes over the
// 2N grid, each intermediate written once and read once, the kernel read once
// per apply for all coils (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:193-218
// runs per coil: fill, pad, batched cuFFT, kernel_mul!, cuFFT, crop).
// T1 and T2 hold ComplexF32, ComplexF64 or (the f64c mode) cx64 values.
//   A  rows (y<Ny, z<Nz) of S_c b_k, zero-padded to 2Nx, FFT along x  -> T1[c][k][z][y][qx]
//   B  columns along y, zero-padded to 2Ny, FFT                       -> T2[c][k][z][qy][qx]
//   C  lines along z: pad to 2Nz, FFT, K x K kernel multiply, inverse FFT, crop to Nz (in place)
//   D  columns along y: inverse FFT, crop to Ny                       -> T1
//   E  rows: inverse FFT along x, crop to Nx, conj(S_c) multiply, coil sum -> partial
#include "apply3d_c.hpp"
#include "fftlines.hpp"

#include <cstdlib>
#include <type_traits>

namespace msr {

template <int NX2, class CT>
constexpr int rows3_block() {
    constexpr int R = sizeof(CT) == 16 ? 8 : 16;  // rows per workgroup (LDS)
    return (R * FFTPlan<NX2>::T) < 256 ? R * FFTPlan<NX2>::T : 256;
}

template <int NX2, class CT, class ST>
__global__ void __launch_bounds__(256)
pass_a_3d(const CT* __restrict__ b, const float2* __restrict__ cmaps, ST* __restrict__ T1,
          const CT* __restrict__ tw, int nx, int64_t nrows, int K, const int* __restrict__ skip,
          int pair) {
    if (skip && *skip) return;
    constexpr int TP = FFTPlan<NX2>::T;
    constexpr int BLK = rows3_block<NX2, CT>();
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NX2, CT>() + 1;
    __shared__ CT s_tw[NX2];
    __shared__ CT s[F * LD];
    load_twiddles<NX2>(s_tw, tw);
    const int64_t r0 = (int64_t)blockIdx.x * F;
    const int k = blockIdx.y, c = blockIdx.z;
    const int64_t V = nrows * nx;
    const CT* bk = b + (int64_t)k * V;
    const float2* cm = cmaps ? cmaps + (int64_t)c * V : nullptr;
    for (int e = threadIdx.x; e < F * NX2; e += BLK) {
        const int l = e / NX2, xx = e - l * NX2;
        const int64_t row = r0 + l;
        CT v = cz<CT>();
        if (row < nrows && xx < nx) {
            v = bk[row * nx + xx];
            if (cm) v = cmul(cto<CT>(cm[row * nx + xx]), v);
        }
        s[l * LD + lds_pad<NX2, CT>(xx)] = v;
    }
    __syncthreads();
    const int g = threadIdx.x / TP, t = threadIdx.x - g * TP;
    fft_lds<NX2, -1>(s + g * LD, s_tw, t);
    ST* out = T1 + ((int64_t)c * K + k) * nrows * NX2;
    if constexpr (std::is_same<ST, cx64>::value) {
        if (pair) {
            // two consecutive cx64 values per 16-byte store (NX2 even)
            for (int e = threadIdx.x; e < F * (NX2 / 2); e += BLK) {
                const int l = e / (NX2 / 2), q = 2 * (e - l * (NX2 / 2));
                const int64_t row = r0 + l;
                if (row < nrows) {
                    const CT v0 = s[l * LD + lds_pad<NX2, CT>(q)], v1 = s[l * LD + lds_pad<NX2, CT>(q + 1)];
                    const cx64 a0 = cx64_encode(v0.x, v0.y), a1 = cx64_encode(v1.x, v1.y);
                    *reinterpret_cast<uint4*>(out + row * NX2 + q) = make_uint4(a0.w0, a0.w1, a1.w0, a1.w1);
                }
            }
            return;
        }
    }
    for (int e = threadIdx.x; e < F * NX2; e += BLK) {
        const int l = e / NX2, q = e - l * NX2;
        const int64_t row = r0 + l;
        if (row < nrows) st_store<CT, ST>(out + row * NX2 + q, s[l * LD + lds_pad<NX2, CT>(q)]);
    }
}

template <int NX2, class CT, class ST>
__global__ void __launch_bounds__(256)
pass_e_3d(const ST* __restrict__ T1, const float2* __restrict__ cmaps,
          CT* __restrict__ partial, const CT* __restrict__ tw, int nx, int64_t nrows,
          int K, int C, int G, const int* __restrict__ skip, int pair) {
    if (skip && *skip) return;
    constexpr int TP = FFTPlan<NX2>::T;
    constexpr int BLK = rows3_block<NX2, CT>();
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NX2, CT>() + 1;
    constexpr int NXC = NX2 / 2;
    constexpr int EA = (F * NXC + BLK - 1) / BLK;
    __shared__ CT s_tw[NX2];
    __shared__ CT s[F * LD];
    load_twiddles<NX2>(s_tw, tw);
    const int64_t r0 = (int64_t)blockIdx.x * F;
    const int k = blockIdx.y, grp = blockIdx.z;
    const int c0 = (int)((int64_t)C * grp / G), c1 = (int)((int64_t)C * (grp + 1) / G);
    const int64_t V = nrows * nx;
    CT acc[EA];
#pragma unroll
    for (int i = 0; i < EA; ++i) acc[i] = cz<CT>();
    const int g = threadIdx.x / TP, t = threadIdx.x - g * TP;
    for (int c = c0; c < c1; ++c) {
        const ST* in = T1 + ((int64_t)c * K + k) * nrows * NX2;
        bool paired = false;
        if constexpr (std::is_same<ST, cx64>::value) {
            if (pair) {
                // two consecutive cx64 values per 16-byte load (NX2 even)
                for (int e = threadIdx.x; e < F * (NX2 / 2); e += BLK) {
                    const int l = e / (NX2 / 2), q = 2 * (e - l * (NX2 / 2));
                    const int64_t row = r0 + l;
                    CT v0 = cz<CT>(), v1 = cz<CT>();
                    if (row < nrows) {
                        const uint4 w = *reinterpret_cast<const uint4*>(in + row * NX2 + q);
                        v0 = cx64_decode(cx64{w.x, w.y});
                        v1 = cx64_decode(cx64{w.z, w.w});
                    }
                    s[l * LD + lds_pad<NX2, CT>(q)] = v0;
                    s[l * LD + lds_pad<NX2, CT>(q + 1)] = v1;
                }
                paired = true;
            }
        }
        if (!paired) {
            for (int e = threadIdx.x; e < F * NX2; e += BLK) {
                const int l = e / NX2, q = e - l * NX2;
                const int64_t row = r0 + l;
                s[l * LD + lds_pad<NX2, CT>(q)] = row < nrows ? st_load<CT, ST>(in + row * NX2 + q) : cz<CT>();
            }
        }
        __syncthreads();
        fft_lds<NX2, +1>(s + g * LD, s_tw, t);
        const float2* cm = cmaps ? cmaps + (int64_t)c * V : nullptr;
#pragma unroll
        for (int i = 0; i < EA; ++i) {
            const int e = threadIdx.x + i * BLK;
            if (e < F * NXC) {
                const int l = e / NXC, xx = e - l * NXC;
                const int64_t row = r0 + l;
                if (row < nrows) {
                    const CT v = s[l * LD + lds_pad<NX2, CT>(xx)];
                    if (cm) cfmac(acc[i], cto<CT>(cm[row * nx + xx]), v);
                    else acc[i] = cadd(acc[i], v);
                }
            }
        }
        __syncthreads();
    }
    CT* pk = partial + ((int64_t)grp * K + k) * V;
#pragma unroll
    for (int i = 0; i < EA; ++i) {
        const int e = threadIdx.x + i * BLK;
        if (e < F * NXC) {
            const int l = e / NXC, xx = e - l * NXC;
            const int64_t row = r0 + l;
            if (row < nrows) pk[row * nx + xx] = acc[i];
        }
    }
}

static void launch_pass_c_3d(const Apply3DArgs& a, hipStream_t st, const int* skip) {
    switch (a.K) {
        case 1: launch_pass_c_3d_k1(a, st, skip); break;
        case 2: launch_pass_c_3d_k2(a, st, skip); break;
        case 3: launch_pass_c_3d_k3(a, st, skip); break;
        case 4: launch_pass_c_3d_k4(a, st, skip); break;
        case 5: launch_pass_c_3d_k5(a, st, skip); break;
        case 6: launch_pass_c_3d_k6(a, st, skip); break;
        case 7: launch_pass_c_3d_k7(a, st, skip); break;
        case 8: launch_pass_c_3d_k8(a, st, skip); break;
        default: fail(MSR_ERR_UNSUPPORTED, "number of coefficients K = %d not in 1..8", a.K);
    }
    MSR_CHECK_LAUNCH();
}

// 16-byte accesses of two consecutive cx64 row values in passes A and E: T1
// 16-byte aligned (each row of 2Nx values then is), and MSR_COLS_SINGLE=1 not
// set (one value per access, to compare)
static int rows3_pair(const void* T1) {
    const char* e = std::getenv("MSR_COLS_SINGLE");
    return (e == nullptr || e[0] != '1') && (reinterpret_cast<uintptr_t>(T1) & 15u) == 0 ? 1 : 0;
}

template <class CT, class ST>
static void run_row_passes_a(const Apply3DArgs& a, const void* b, hipStream_t st, const int* skip) {
    const int NX2 = 2 * a.nx;
    const int64_t nrows = (int64_t)a.ny * a.nz;
    bool ok = dispatch_fft_len(NX2, [&]<int N>() {
        if constexpr (N <= 512) {
            constexpr int F = rows3_block<N, CT>() / FFTPlan<N>::T;
            dim3 grid((unsigned)ceil_div(nrows, F), a.K, a.C);
            hipLaunchKernelGGL((pass_a_3d<N, CT, ST>), grid, dim3(rows3_block<N, CT>()), 0, st,
                               static_cast<const CT*>(b), a.cmaps, static_cast<ST*>(a.T1),
                               static_cast<const CT*>(a.tw_x), a.nx, nrows, a.K, skip, rows3_pair(a.T1));
        } else {
            fail(MSR_ERR_UNSUPPORTED, "Toeplitz grid 2Nx = %d above 512", N);
        }
    });
    MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Nx = %d", NX2);
    MSR_CHECK_LAUNCH();
}

template <class CT, class ST>
static void run_row_passes_e(const Apply3DArgs& a, hipStream_t st, const int* skip) {
    const int64_t nrows = (int64_t)a.ny * a.nz;
    dispatch_fft_len(2 * a.nx, [&]<int N>() {
        if constexpr (N <= 512) {
            constexpr int F = rows3_block<N, CT>() / FFTPlan<N>::T;
            dim3 grid((unsigned)ceil_div(nrows, F), a.K, a.G);
            hipLaunchKernelGGL((pass_e_3d<N, CT, ST>), grid, dim3(rows3_block<N, CT>()), 0, st,
                               static_cast<const ST*>(a.T1), a.cmaps, static_cast<CT*>(a.partial),
                               static_cast<const CT*>(a.tw_x), a.nx, nrows, a.K, a.C, a.G, skip,
                               rows3_pair(a.T1));
        }
    });
    MSR_CHECK_LAUNCH();
}

static void run_passes(const Apply3DArgs& a, msr_ctx* ctx, const void* b, hipStream_t st,
                       const int* skip) {
    const int NX2 = 2 * a.nx;
    if (!a.f64) run_row_passes_a<float2, float2>(a, b, st, skip);
    else if (a.compact) run_row_passes_a<double2, cx64>(a, b, st, skip);
    else run_row_passes_a<double2, double2>(a, b, st, skip);
    // B: y columns of T1 [(c,k,z)][y][qx] -> T2 [(c,k,z)][qy][qx]
    LinePass pb;
    pb.in = a.T1;
    pb.out = a.T2;
    pb.f64 = a.f64;
    pb.nf = 2 * a.ny;
    pb.dir = -1;
    pb.lin = a.ny;
    pb.lout = 2 * a.ny;
    pb.sel = SEL_FIRST;
    pb.nb = (int64_t)a.C * a.K * a.nz;
    pb.ncols = NX2;
    pb.in_bstride = (int64_t)a.ny * NX2;
    pb.in_stride = NX2;
    pb.out_bstride = (int64_t)2 * a.ny * NX2;
    pb.out_stride = NX2;
    pb.skip = skip;
    pb.in_compact = pb.out_compact = a.compact;
    run_cols(pb, ctx, st);
    launch_pass_c_3d(a, st, skip);
    // D: inverse along y, crop to Ny: T2 -> T1
    LinePass pd = pb;
    pd.in = a.T2;
    pd.out = a.T1;
    pd.dir = +1;
    pd.lin = 2 * a.ny;
    pd.lout = a.ny;
    pd.in_bstride = (int64_t)2 * a.ny * NX2;
    pd.out_bstride = (int64_t)a.ny * NX2;
    run_cols(pd, ctx, st);
    if (!a.f64) run_row_passes_e<float2, float2>(a, st, skip);
    else if (a.compact) run_row_passes_e<double2, cx64>(a, st, skip);
    else run_row_passes_e<double2, double2>(a, st, skip);
}

static Apply3DArgs args3d(msr_normalop* op, int am) {
    const bool f64 = am != AM_F32;
    Apply3DArgs a;
    a.nx = (int)op->N[0];
    a.ny = (int)op->N[1];
    a.nz = (int)op->N[2];
    a.K = op->K;
    a.C = op->C;
    a.f64 = f64 ? 1 : 0;
    a.compact = am == AM_F64C ? 1 : 0;
    a.cmaps = op->identity_cmaps ? nullptr : op->cmaps.as<float2>();
    a.lam = op->lam->p;
    a.lam_real = op->lam_real;
    a.T1 = op->T.p;
    a.T2 = op->T2.p;
    a.partial = op->partial.p;
    a.G = op->G;
    a.Gb = op->Gb;
    if (f64) {
        a.tw_x = twiddle_table64(op->ctx, 2 * a.nx);
        a.tw_y = twiddle_table64(op->ctx, 2 * a.ny);
        a.tw_z = twiddle_table64(op->ctx, 2 * a.nz);
    } else {
        a.tw_x = twiddle_table(op->ctx, 2 * a.nx);
        a.tw_y = twiddle_table(op->ctx, 2 * a.ny);
        a.tw_z = twiddle_table(op->ctx, 2 * a.nz);
    }
    return a;
}

void launch_apply3d_passes(msr_normalop* op, const void* b, int am, hipStream_t st,
                           const int* skip) {
    run_passes(args3d(op, am), op->ctx, b, st, skip);
}

void launch_apply3d_passes_coil(msr_normalop* op, const float2* b, int coil, hipStream_t st) {
    Apply3DArgs a = args3d(op, AM_F32);
    a.C = 1;
    a.G = 1;
    a.Gb = 1;
    if (a.cmaps) a.cmaps += (size_t)coil * op->V;
    run_passes(a, op->ctx, b, st, nullptr);
}

}  // namespace msr
