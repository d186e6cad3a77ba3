// Pass B of the fused 2D apply: column FFT, per-voxel K x K kernel multiply,
// inverse column FFT (see apply2d.hpp).  Templated on the column length and K;
// the instantiations are spread over apply2d_b_k*.hip to build in parallel.
#pragma once

#include "apply2d.hpp"
#include "fft.hpp"

namespace msr {

// 0-based LAPACK 'U' packed index of (j, k), j <= k
// (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:239, 1-based j + k(k-1)/2).
__host__ __device__ constexpr int packed_index(int j, int k) { return j + k * (k + 1) / 2; }

template <int NY2, int K>
struct PassBShape {
    static constexpr int TP = FFTPlan<NY2>::T;
    static constexpr int WC0 = (256 / (K * TP)) > 1 ? 256 / (K * TP) : 1;
    static constexpr int WC = WC0 < 16 ? WC0 : 16;  // columns per WG
    static constexpr int BLK = WC * K * TP;
    static constexpr int NV = WC * NY2;                // voxels per WG
    static constexpr int EV = (NV + BLK - 1) / BLK;    // voxels per thread
    static constexpr int P = K * (K + 1) / 2;
};

template <int NY2, int K, bool LREAL>
__global__ void __launch_bounds__(512)
pass_b_2d(float2* __restrict__ T, const void* __restrict__ lam_v, const float2* __restrict__ tw,
          int nx2, int ny, int C, int Gb, const int* __restrict__ skip) {
    if (skip && *skip) return;
    using S = PassBShape<NY2, K>;
    using LT = typename std::conditional<LREAL, float, float2>::type;
    constexpr int TP = S::TP, WC = S::WC, BLK = S::BLK, NV = S::NV, EV = S::EV, P = S::P;
    constexpr int LL = lds_line<NY2>();
    __shared__ float2 s_tw[NY2];
    __shared__ float2 s_buf[WC * K * LL];
    const LT* lam = reinterpret_cast<const LT*>(lam_v);
    const int tid = threadIdx.x;
    const int qx0 = blockIdx.x * WC;
    const int grp = blockIdx.y;
    const int c0 = (int)((int64_t)C * grp / Gb), c1 = (int)((int64_t)C * (grp + 1) / Gb);
    load_twiddles<NY2>(s_tw, tw);

    // kernel values of the voxels this thread owns, kept for all coils
    LT lr[EV][P];
#pragma unroll
    for (int e = 0; e < EV; ++e) {
        const int v = tid + e * BLK;
        const int wc = v / NY2, qy = v - wc * NY2;
        const int qx = qx0 + wc;
        const bool ok = (v < NV) && (qx < nx2);
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (ok) lr[e][p] = lam[((size_t)qx * P + p) * NY2 + qy];
            else if constexpr (LREAL) lr[e][p] = 0.f;
            else lr[e][p] = make_float2(0.f, 0.f);
        }
    }

    const int g = tid / TP, t = tid - g * TP;
    for (int c = c0; c < c1; ++c) {
        // load K columns per owned qx, zero-padded from ny to NY2
        for (int el = tid; el < WC * K * NY2; el += BLK) {
            const int line = el / NY2, i = el - line * NY2;
            const int wc = line / K, k = line - wc * K;
            const int qx = qx0 + wc;
            float2 v = make_float2(0.f, 0.f);
            if (i < ny && qx < nx2) v = T[(((size_t)c * K + k) * nx2 + qx) * ny + i];
            s_buf[line * LL + lds_pad(i)] = v;
        }
        __syncthreads();
        fft_lds<NY2, -1>(s_buf + g * LL, s_tw, t);
        // per-voxel Hermitian (or real symmetric) K x K multiply
#pragma unroll
        for (int e = 0; e < EV; ++e) {
            const int v = tid + e * BLK;
            if (v < NV) {
                const int wc = v / NY2, qy = v - wc * NY2;
                float2* col = s_buf + (size_t)wc * K * LL + lds_pad(qy);
                float2 in[K];
#pragma unroll
                for (int k = 0; k < K; ++k) in[k] = col[k * LL];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        if constexpr (LREAL) {
                            const float l = lr[e][j <= k ? packed_index(j, k) : packed_index(k, j)];
                            acc.x = fmaf(l, in[k].x, acc.x);
                            acc.y = fmaf(l, in[k].y, acc.y);
                        } else {
                            if (j <= k) cfma(acc, lr[e][packed_index(j, k)], in[k]);
                            else cfmac(acc, lr[e][packed_index(k, j)], in[k]);
                        }
                    }
                    col[j * LL] = acc;
                }
            }
        }
        __syncthreads();
        fft_lds<NY2, +1>(s_buf + g * LL, s_tw, t);
        for (int el = tid; el < WC * K * NY2; el += BLK) {
            const int line = el / NY2, i = el - line * NY2;
            const int wc = line / K, k = line - wc * K;
            const int qx = qx0 + wc;
            if (i < ny && qx < nx2)
                T[(((size_t)c * K + k) * nx2 + qx) * ny + i] = s_buf[line * LL + lds_pad(i)];
        }
        __syncthreads();
    }
}

template <int NY2, int K, bool LREAL>
void launch_pass_b_2d_t(const Apply2DArgs& a, hipStream_t st, const int* skip) {
    using S = PassBShape<NY2, K>;
    const int nx2 = 2 * a.nx;
    dim3 grid((nx2 + S::WC - 1) / S::WC, a.Gb);
    hipLaunchKernelGGL((pass_b_2d<NY2, K, LREAL>), grid, dim3(S::BLK), 0, st, a.T, a.lam, a.tw_y,
                       nx2, a.ny, a.C, a.Gb, skip);
}

// per-K dispatchers (defined in apply2d_b_k*.hip)
#define MSR_DECL_PASSB(KK) void launch_pass_b_2d_k##KK(const Apply2DArgs& a, hipStream_t st, const int* skip);
MSR_DECL_PASSB(1) MSR_DECL_PASSB(2) MSR_DECL_PASSB(3) MSR_DECL_PASSB(4)
MSR_DECL_PASSB(5) MSR_DECL_PASSB(6) MSR_DECL_PASSB(7) MSR_DECL_PASSB(8)
#undef MSR_DECL_PASSB

inline void launch_pass_b_2d(const Apply2DArgs& a, hipStream_t st, const int* skip) {
    switch (a.K) {
        case 1: launch_pass_b_2d_k1(a, st, skip); break;
        case 2: launch_pass_b_2d_k2(a, st, skip); break;
        case 3: launch_pass_b_2d_k3(a, st, skip); break;
        case 4: launch_pass_b_2d_k4(a, st, skip); break;
        case 5: launch_pass_b_2d_k5(a, st, skip); break;
        case 6: launch_pass_b_2d_k6(a, st, skip); break;
        case 7: launch_pass_b_2d_k7(a, st, skip); break;
        case 8: launch_pass_b_2d_k8(a, st, skip); break;
        default: fail(MSR_ERR_UNSUPPORTED, "number of coefficients K = %d not in 1..8", a.K);
    }
    MSR_CHECK_LAUNCH();
}

// Body of a per-K dispatcher: switch over the supported column lengths.
#define MSR_DEFINE_PASSB(KK)                                                                   \
    void launch_pass_b_2d_k##KK(const Apply2DArgs& a, hipStream_t st, const int* skip) {       \
        const int ny2 = 2 * a.ny;                                                              \
        bool ok = dispatch_fft_len(ny2, [&]<int N>() {                                         \
            if constexpr (N <= 512) {                                                          \
                if (a.lam_real) launch_pass_b_2d_t<N, KK, true>(a, st, skip);                  \
                else launch_pass_b_2d_t<N, KK, false>(a, st, skip);                            \
            } else {                                                                           \
                fail(MSR_ERR_UNSUPPORTED, "Toeplitz grid 2Ny = %d above 512", N);             \
            }                                                                                  \
        });                                                                                    \
        MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Ny = %d", ny2);           \
    }

}  // namespace msr
