// Pass B of the fused 2D apply: column FFT, per-voxel K x K kernel multiply,
// inverse column FFT (see apply2d.hpp).  Templated on the column length and K;
// the instantiations are spread over apply2d_b_k*.hip to build in parallel.
#pragma once

#include "apply2d.hpp"
#include "fft.hpp"

namespace msr {
#ifndef PB_WAVES
#define PB_WAVES 4
#endif
#ifndef PB_LAMG
#define PB_LAMG false
#endif
#ifndef PB_PREFETCH
#define PB_PREFETCH false
#endif
#ifndef PB_CROPREG
#define PB_CROPREG true
#endif

// 0-based LAPACK 'U' packed index of (j, k), j <= k
// (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:239, 1-based j + k(k-1)/2).
__host__ __device__ constexpr int packed_index(int j, int k) { return j + k * (k + 1) / 2; }

// acc[j] += sum_k Lambda_jk in[k] for the packed Hermitian (or real
// symmetric) K x K kernel of one voxel, lr[packed_index(j, k)] for j <= k and
// Lambda_kj = conj(Lambda_jk).  Each packed entry is promoted once, on use, to
// the arithmetic type, so the registers hold the Float32 table only.
template <int K, bool LREAL, class CT, class LT>
__device__ __forceinline__ void kernel_multiply(const LT* lr, const CT* in, CT* acc) {
    using RT = real_t<CT>;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j <= k; ++j) {
            const LT lf = sizeof(CT) == 16 ? opaque(lr[packed_index(j, k)]) : lr[packed_index(j, k)];
            if constexpr (LREAL) {
                const RT l = (RT)lf;
                acc[j].x = fma(l, in[k].x, acc[j].x);
                acc[j].y = fma(l, in[k].y, acc[j].y);
                if (j != k) {
                    acc[k].x = fma(l, in[j].x, acc[k].x);
                    acc[k].y = fma(l, in[j].y, acc[k].y);
                }
            } else {
                const CT l = cto<CT>(lf);
                cfma(acc[j], l, in[k]);
                if (j != k) cfmac(acc[k], l, in[j]);
            }
        }
}

// Workgroup shape: WC grid columns, each with K lines of NY2 (FFT by
// TP = FFTPlan<NY2>::T threads per line), and at least one thread per voxel of
// the WC columns, so every thread holds the kernel of one voxel in registers
// (threads beyond the K*WC*TP line threads idle during the FFT stages).
template <int NY2, int K, class CT>
struct PassBShape {
    static constexpr int TP = FFTPlan<NY2>::T;
    static constexpr int PER = (K * TP > NY2) ? K * TP : NY2;  // threads per column
    static constexpr int WC0 = (512 / PER) > 1 ? 512 / PER : 1;
    static constexpr int WCM = sizeof(CT) == 16 ? 8 : 16;  // LDS: 8 ComplexF64 columns
    static constexpr int WC = WC0 < WCM ? WC0 : WCM;        // columns per WG
    static constexpr int FT = WC * K * TP;                  // FFT threads
    static constexpr int BLK = WC * PER;
    static constexpr int NV = WC * NY2;                // voxels per WG
    static constexpr int EV = (NV + BLK - 1) / BLK;    // voxels per thread (1)
    static constexpr int P = K * (K + 1) / 2;
};

// block lb = (column block, coil group), column block fastest.  The kernel
// stays ComplexF32 (or Float32) and is promoted to the arithmetic type CT; the
// intermediate T is stored as ST.
template <int NY2, int K, bool LREAL, class CT, class ST>
__device__ __forceinline__ void pass_b_body(const Apply2DArgs& a, int64_t lb) {
    if (a.skip && *a.skip) return;
    using S = PassBShape<NY2, K, CT>;
    using LT = typename std::conditional<LREAL, float, float2>::type;
    using RT = real_t<CT>;
    constexpr int TP = S::TP, WC = S::WC, BLK = S::BLK, NV = S::NV, EV = S::EV, P = S::P;
    constexpr int FT = S::FT;
    // the swizzled LDS layout up to K = 5 (above, its address arithmetic
    // makes the pass spill)
    constexpr bool SWZ = lds_swizzled<NY2, CT>() && K <= 5;
    constexpr int LL = lds_line<NY2, CT, SWZ>();
    // ComplexF64 twiddles are read from the global table (L1 / L2), which frees
    // LDS for one more workgroup per CU; ComplexF32 ones from an LDS copy
    __shared__ CT s_tw_lds[sizeof(CT) == 16 ? 1 : NY2];
    const CT* s_tw = sizeof(CT) == 16 ? static_cast<const CT*>(a.tw_y) : s_tw_lds;
    __shared__ CT s_buf[WC * K * LL];
    const LT* lam = reinterpret_cast<const LT*>(a.lam);
    ST* T = static_cast<ST*>(a.T);
    const int nx2 = 2 * a.nx, ny = a.ny, C = a.C, Gb = a.Gb;
    const int ncb = (nx2 + WC - 1) / WC;
    const int tid = threadIdx.x;
    const int qx0 = (int)(lb % ncb) * WC;
    const int grp = (int)(lb / ncb);
    const int c0 = (int)((int64_t)C * grp / Gb), c1 = (int)((int64_t)C * (grp + 1) / Gb);
    if constexpr (sizeof(CT) != 16)
        load_twiddles<NY2>(s_tw_lds, static_cast<const CT*>(a.tw_y));

    // kernel values of the voxels this thread owns: kept in registers for all
    // coils, or (LAMG) read again from global memory for every coil
    constexpr bool LAMG = PB_LAMG && sizeof(CT) == 16;
    auto load_lam = [&](LT* dst, int e) {
        const int v = tid + e * BLK;
        const int wc = v / NY2, qy = v - wc * NY2;
        const int qx = qx0 + wc;
        const bool ok = (v < NV) && (qx < nx2);
        // branch-free: an out-of-range voxel reads column 0 and keeps zero
        const LT* src = lam + (ok ? (size_t)qx * P * NY2 + qy : 0);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const LT x = src[(size_t)p * NY2];
            if constexpr (LREAL) dst[p] = ok ? x : 0.f;
            else dst[p] = ok ? x : make_float2(0.f, 0.f);
        }
    };
    LT lr[LAMG ? 1 : EV][P];
    if constexpr (!LAMG) {
#pragma unroll
        for (int e = 0; e < EV; ++e) load_lam(lr[e], e);
    }

    // FFT thread (g, t): line g = wc * K + k of the WG, t < TP.  Its line is
    // held in registers in the transform's first-stage layout, element
    // i = t + p TP + r NY2/R0 at reg[p R0 + r]; the elements of the upper half
    // (i >= ny, the zero padding) are the constant zero.
    constexpr int NR = NY2 / TP;
    constexpr bool PREFETCH = PB_PREFETCH;
    constexpr int R0 = FFTPlan<NY2>::R[0], NB0 = NY2 / R0;
    const int g = tid / TP, t = tid - g * TP;
    const bool fft_thread = tid < FT;
    const int gl = fft_thread ? g : 0;
    const int wcl = gl / K, kl = gl - wcl * K;
    const bool col_ok = fft_thread && (qx0 + wcl) < nx2;
    CT* line = s_buf + gl * LL;
    CT pre[NR];  // the next coil's line, loaded one coil ahead
    auto prefetch = [&](int c) {
        // branch-free: a thread without a line reads line 0 of the WG and keeps zero
        const ST* src = T + (((size_t)c * K + kl) * nx2 + (col_ok ? qx0 + wcl : qx0)) * ny;
#pragma unroll
        for (int p = 0; p < NR / R0; ++p)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const int i = t + p * TP + r * NB0;
                if (r * NB0 < NY2 / 2) {
                    const CT x = st_load<CT, ST>(src + (i < NY2 / 2 ? i : 0));
                    pre[p * R0 + r] = (i < NY2 / 2 && col_ok) ? x : cz<CT>();
                } else {
                    pre[p * R0 + r] = cz<CT>();
                }
            }
    };
    if (PREFETCH && c0 < c1) prefetch(c0);
    __syncthreads();  // twiddles
    for (int c = c0; c < c1; ++c) {
        CT xr[NR];
        if constexpr (!PREFETCH) prefetch(c);
#pragma unroll
        for (int e = 0; e < NR; ++e) xr[e] = pre[e];
        if (PREFETCH && c + 1 < c1) prefetch(c + 1);  // in flight during this coil's transforms
        if (fft_thread) fft_wave<NY2, -1, true, false, CT, SWZ>(line, s_tw, t, true, xr);
        __syncthreads();
        // per-voxel Hermitian (or real symmetric) K x K multiply
#pragma unroll
        for (int e = 0; e < EV; ++e) {
            const int v = tid + e * BLK;
            if (v < NV) {
                const int wc = v / NY2, qy = v - wc * NY2;
                CT* col = s_buf + (size_t)wc * K * LL + lds_pad<NY2, CT, SWZ>(qy);
                CT in[K], acc[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    in[k] = col[k * LL];
                    acc[k] = cz<CT>();
                }
                if constexpr (LAMG) {
                    LT lg[P];
                    load_lam(lg, e);
                    kernel_multiply<K, LREAL, CT>(lg, in, acc);
                } else {
                    kernel_multiply<K, LREAL, CT>(lr[e], in, acc);
                }
#pragma unroll
                for (int j = 0; j < K; ++j) col[j * LL] = acc[j];
            }
        }
        __syncthreads();
        if (fft_thread) {
            ST* dst = T + (((size_t)c * K + kl) * nx2 + qx0 + wcl) * ny;
            constexpr int RL = FFTPlan<NY2>::R[FFTPlan<NY2>::NST - 1], NBL = NY2 / RL;
            if constexpr (PB_CROPREG && RL % 2 == 0 && sizeof(ST) <= 8) {
                // the last stage leaves X[t + p TP + r NY2/RL] in xr[p RL + r]; the
                // crop keeps i < NY2/2, i.e. r < RL/2 (the other outputs are dead),
                // stored straight from registers, lanes on consecutive i.  For 8-byte
                // storage only: at 256^2 / 20 coils / K = 5 on MI355X it took the
                // apply from 0.219 to 0.212 ms (ComplexF32) and 0.289 to 0.277 ms
                // (compact), but ComplexF64 storage from 0.302 to 0.325 ms
                fft_wave<NY2, +1, false, true, CT, SWZ>(line, s_tw, t, true, xr);
#pragma unroll
                for (int p = 0; p < NBL / TP; ++p)
#pragma unroll
                    for (int r = 0; r < RL / 2; ++r)
                        if (col_ok) st_store<CT, ST>(dst + t + p * TP + r * NBL, xr[p * RL + r]);
            } else {
                // the last stage writes the line to LDS; the crop stores
                // elements i < ny, lanes on consecutive i
                fft_wave<NY2, +1, false, false, CT, SWZ>(line, s_tw, t, true, xr);
#pragma unroll
                for (int m = 0; m < NR / 2; ++m) {
                    const int i = t + m * TP;
                    if (col_ok) st_store<CT, ST>(dst + i, line[lds_pad<NY2, CT, SWZ>(i)]);
                }
            }
        }
    }
}

template <int NY2, int K, bool LREAL, class CT, class ST>
__global__ void __launch_bounds__(512, PB_WAVES) pass_b_2d(Apply2DArgs a) {
    pass_b_body<NY2, K, LREAL, CT, ST>(a, blockIdx.x);
}

template <int NY2, int K, bool LREAL, class CT, class ST>
__global__ void __launch_bounds__(512, PB_WAVES)
pass_b_2d_batch(const Apply2DArgs* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    pass_b_body<NY2, K, LREAL, CT, ST>(jobs[j], (int64_t)blockIdx.x - pre[j]);
}

template <int NY2, int K, bool LREAL, class CT, class ST>
void launch_pass_b_t(const Apply2DArgs& a, const Apply2DArgs* jobs, const int* pre, int nj,
                     int64_t nblocks, hipStream_t st) {
    using S = PassBShape<NY2, K, CT>;
    dim3 grid((unsigned)nblocks);
    if (jobs)
        hipLaunchKernelGGL((pass_b_2d_batch<NY2, K, LREAL, CT, ST>), grid, dim3(S::BLK), 0, st,
                           jobs, pre, nj);
    else
        hipLaunchKernelGGL((pass_b_2d<NY2, K, LREAL, CT, ST>), grid, dim3(S::BLK), 0, st, a);
}

// the arithmetic / storage pair of the apply mode
template <int NY2, int K, bool LREAL>
void launch_pass_b_mode(const Apply2DArgs& a, const Apply2DArgs* jobs, const int* pre, int nj,
                        int64_t nblocks, hipStream_t st) {
    if (!a.f64) launch_pass_b_t<NY2, K, LREAL, float2, float2>(a, jobs, pre, nj, nblocks, st);
    else if (a.compact) launch_pass_b_t<NY2, K, LREAL, double2, cx64>(a, jobs, pre, nj, nblocks, st);
    else launch_pass_b_t<NY2, K, LREAL, double2, double2>(a, jobs, pre, nj, nblocks, st);
}

// per-K entry points (defined in apply2d_b_k*.hip); jobs == nullptr means the
// single-operator launch of `a`
#define MSR_DECL_PASSB(KK)                                                                     \
    void pass_b_2d_k##KK(const Apply2DArgs& a, const Apply2DArgs* jobs, const int* pre, int nj, \
                         int64_t nblocks, hipStream_t st);                                     \
    int pass_b_2d_colblocks_k##KK(int nx2, int ny2, int f64);
MSR_DECL_PASSB(1) MSR_DECL_PASSB(2) MSR_DECL_PASSB(3) MSR_DECL_PASSB(4)
MSR_DECL_PASSB(5) MSR_DECL_PASSB(6) MSR_DECL_PASSB(7) MSR_DECL_PASSB(8)
#undef MSR_DECL_PASSB

#define MSR_PASSB_SWITCH(K, CALL)                                                              \
    switch (K) {                                                                               \
        case 1: CALL(1); break;                                                                \
        case 2: CALL(2); break;                                                                \
        case 3: CALL(3); break;                                                                \
        case 4: CALL(4); break;                                                                \
        case 5: CALL(5); break;                                                                \
        case 6: CALL(6); break;                                                                \
        case 7: CALL(7); break;                                                                \
        case 8: CALL(8); break;                                                                \
        default: fail(MSR_ERR_UNSUPPORTED, "number of coefficients K = %d not in 1..8", K);   \
    }

inline void launch_pass_b_2d(const Apply2DArgs& a, hipStream_t st) {
#define MSR_CALLB(KK) pass_b_2d_k##KK(a, nullptr, nullptr, 0, blocks_b2d(a), st)
    MSR_PASSB_SWITCH(a.K, MSR_CALLB)
#undef MSR_CALLB
    MSR_CHECK_LAUNCH();
}

// Body of a per-K entry point: switch over the supported column lengths.
#define MSR_DEFINE_PASSB(KK)                                                                   \
    void pass_b_2d_k##KK(const Apply2DArgs& a, const Apply2DArgs* jobs, const int* pre, int nj, \
                         int64_t nblocks, hipStream_t st) {                                    \
        const int ny2 = 2 * a.ny;                                                              \
        MSR_REQUIRE(nblocks >= 1 && nblocks < ((int64_t)1 << 31), MSR_ERR_ARGUMENT,            \
                    "pass B launch of %lld blocks", (long long)nblocks);                       \
        bool ok = dispatch_fft_len(ny2, [&]<int N>() {                                         \
            if constexpr (N <= 512) {                                                          \
                if (a.lam_real) launch_pass_b_mode<N, KK, true>(a, jobs, pre, nj, nblocks, st); \
                else launch_pass_b_mode<N, KK, false>(a, jobs, pre, nj, nblocks, st);         \
            } else {                                                                           \
                fail(MSR_ERR_UNSUPPORTED, "Toeplitz grid 2Ny = %d above 512", N);             \
            }                                                                                  \
        });                                                                                    \
        MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Ny = %d", ny2);           \
    }                                                                                          \
    int pass_b_2d_colblocks_k##KK(int nx2, int ny2, int f64) {                                 \
        int r = -1;                                                                            \
        dispatch_fft_len(ny2, [&]<int N>() {                                                   \
            constexpr int WC = PassBShape<N, KK, float2>::WC;                                  \
            constexpr int WC64 = PassBShape<N, KK, double2>::WC;                               \
            r = (nx2 + (f64 ? WC64 : WC) - 1) / (f64 ? WC64 : WC);                             \
        });                                                                                    \
        MSR_REQUIRE(r > 0, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Ny = %d", ny2);        \
        return r;                                                                              \
    }

}  // namespace msr
