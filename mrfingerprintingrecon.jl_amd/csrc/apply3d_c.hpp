// Pass C of the fused 3D apply: for WC adjacent grid columns (qx) at one qy,
// the K lines along z of every coil are zero-padded from Nz to 2Nz, FFT'd,
// multiplied voxel by voxel by the packed K x K kernel (held in registers for
// the whole coil loop), inverse FFT'd and cropped back to Nz, in place in
//   T2[c][k][z][qy][qx]   (lines strided by 2Nx*2Ny; adjacent qx coalesce).
// Dense kernel layout: lam[((qy * 2Nx + qx) * P + p) * 2Nz + qz].
#pragma once

#include "apply3d.hpp"
#include "apply2d_b.hpp"
#include "fft.hpp"

namespace msr {

// Intermediates and arithmetic in ComplexF32 (f64 = 0) or ComplexF64 (f64 =
// 1); the kernel and coil maps stay ComplexF32.
struct Apply3DArgs {
    int nx, ny, nz, K, C;
    int f64;
    int compact;          // T1, T2 stored as cx64 (with f64)
    const float2* cmaps;  // (C, nz, ny, nx) or nullptr
    const void* lam;
    int lam_real;
    void* T1;             // (C, K, nz, ny, 2nx)
    void* T2;             // (C, K, nz, 2ny, 2nx)
    void* partial;        // (G, K, V)
    int G, Gb;
    const void *tw_x, *tw_y, *tw_z;
};

// WC adjacent qx columns per workgroup, K z lines each (FFT by TP threads per
// line), at least one thread per voxel (the kernel of one voxel per thread in
// registers; threads beyond the K*WC*TP line threads idle in the FFT stages).
template <int NZ2, int K, class CT>
struct PassC3Shape {
    static constexpr int TP = FFTPlan<NZ2>::T;
    static constexpr int PER = (K * TP > NZ2) ? K * TP : NZ2;                // threads per column
    static constexpr int WCT = (512 / PER) > 1 ? 512 / PER : 1;              // thread limit
    static constexpr int LB = 87376 / (int)sizeof(CT);                        // <= ~87 KB LDS
    static constexpr int WCL = (LB / (K * NZ2)) > 1 ? LB / (K * NZ2) : 1;
    static constexpr int WC0 = WCT < WCL ? WCT : WCL;
    static constexpr int WC = WC0 < 8 ? WC0 : 8;     // adjacent qx lines per WG
    static constexpr int FT = WC * K * TP;           // FFT threads
    static constexpr int BLK = WC * PER;
    static constexpr int NV = WC * NZ2;
    static constexpr int EV = (NV + BLK - 1) / BLK;  // voxels per thread (1)
    static constexpr int P = K * (K + 1) / 2;
};

// The register form (below) for exact plans of at most 256 points whose
// threads hold at most 8 values, and K <= 4: bounded to 128 VGPRs (four waves
// per SIMD, two 512-thread workgroups per CU), it spills with more values per
// thread or a larger kernel table.
template <int NZ2, int K>
constexpr bool pc3_register_form() {
    return fft_plan_exact<NZ2>() && NZ2 <= 256 && NZ2 / FFTPlan<NZ2>::T <= 8 && K <= 4;
}

template <int NZ2, int K, bool LREAL, class CT, class ST>
__global__ void __launch_bounds__(512, (pc3_register_form<NZ2, K>() ? 4 : 2))
pass_c_3d(ST* __restrict__ T2, const void* __restrict__ lam_v, const CT* __restrict__ tw,
          int nx2, int ny2, int nz, int C, int Gb, const int* __restrict__ skip) {
    if (skip && *skip) return;
    using S = PassC3Shape<NZ2, K, CT>;
    using LT = typename std::conditional<LREAL, float, float2>::type;
    using RT = real_t<CT>;
    constexpr int TP = S::TP, WC = S::WC, BLK = S::BLK, NV = S::NV, EV = S::EV, P = S::P;
    constexpr int FT = S::FT;
    constexpr int LL = lds_line<NZ2, CT>();
    __shared__ CT s_tw[NZ2];
    __shared__ CT s_buf[WC * K * LL];
    const LT* lam = reinterpret_cast<const LT*>(lam_v);
    const int tid = threadIdx.x;
    // XCD-aware order: workgroups dispatched to one XCD (every 8th) take
    // neighbouring column blocks, so partially used cache lines of the strided
    // z lines are shared in that XCD's L2
    const int nbx = gridDim.x, nxy = gridDim.x * gridDim.y;
    const int lin = blockIdx.y * nbx + blockIdx.x;
    const int per = (nxy + 7) / 8;
    int logical = (lin % 8) * per + lin / 8;
    if ((nxy % 8) != 0) logical = lin;  // keep the plain order when it does not divide
    const int qx0 = (logical % nbx) * WC;
    const int qy = logical / nbx;
    const int grp = blockIdx.z;
    const int c0 = (int)((int64_t)C * grp / Gb), c1 = (int)((int64_t)C * (grp + 1) / Gb);
    const int64_t plane = (int64_t)nx2 * ny2;  // stride between z lines
    load_twiddles<NZ2>(s_tw, tw);

    LT lr[EV][P];
#pragma unroll
    for (int e = 0; e < EV; ++e) {
        const int v = tid + e * BLK;
        const int wc = v / NZ2, qz = v - wc * NZ2;
        const int qx = qx0 + wc;
        const bool ok = (v < NV) && (qx < nx2);
        if constexpr (pc3_register_form<NZ2, K>()) {
            // branch-free, so the P loads issue together: an out-of-range
            // voxel reads column 0 and keeps zero
            const LT* src = lam + (ok ? ((int64_t)qy * nx2 + qx) * P * NZ2 + qz : 0);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const LT x = src[(int64_t)p * NZ2];
                if constexpr (LREAL) lr[e][p] = ok ? x : 0.f;
                else lr[e][p] = ok ? x : make_float2(0.f, 0.f);
            }
        } else {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                if (ok) lr[e][p] = lam[(((int64_t)qy * nx2 + qx) * P + p) * NZ2 + qz];
                else if constexpr (LREAL) lr[e][p] = 0.f;
                else lr[e][p] = make_float2(0.f, 0.f);
            }
        }
    }
    // register form: per-voxel Hermitian (or real symmetric) K x K multiply of
    // the lines in LDS, line (k, wc) at col_base(wc) + k * kstride; the kernel
    // entry is promoted on use (opaque: the registers keep the Float32 table
    // across coils instead of a Float64 copy)
    auto multiply = [&](int kstride, auto col_base) {
#pragma unroll
        for (int e = 0; e < EV; ++e) {
            const int v = tid + e * BLK;
            if (v < NV) {
                const int wc = v / NZ2, qz = v - wc * NZ2;
                CT* col = col_base(wc) + lds_pad<NZ2, CT>(qz);
                CT in[K];
#pragma unroll
                for (int k = 0; k < K; ++k) in[k] = col[k * kstride];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    CT acc = cz<CT>();
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const LT lv = lr[e][j <= k ? packed_index(j, k) : packed_index(k, j)];
                        const LT lf = sizeof(CT) == 16 ? opaque(lv) : lv;
                        if constexpr (LREAL) {
                            const RT l = (RT)lf;
                            acc.x = fma(l, in[k].x, acc.x);
                            acc.y = fma(l, in[k].y, acc.y);
                        } else {
                            if (j <= k) cfma(acc, cto<CT>(lf), in[k]);
                            else cfmac(acc, cto<CT>(lf), in[k]);
                        }
                    }
                    col[j * kstride] = acc;
                }
            }
        }
    };

    if constexpr (pc3_register_form<NZ2, K>()) {
        // Register form (exact plans): line (k, wc) = g = k WC + wc, its TP
        // threads in one wavefront (adjacent columns share a wavefront, so its
        // accesses pair up like the staged form's).  A thread loads its line's
        // elements straight into the first stage's layout (the padding half
        // the constant zero), the forward transform leaves the line in LDS, the
        // multiply runs between two barriers, and the inverse transform leaves
        // the cropped outputs in registers, stored from there: two workgroup
        // barriers per coil, no LDS staging of the stored values.
        constexpr int R0 = FFTPlan<NZ2>::R[0], NB0 = NZ2 / R0;
        constexpr int RL = FFTPlan<NZ2>::R[FFTPlan<NZ2>::NST - 1], NBL = NZ2 / RL;
        constexpr int NR = NZ2 / TP;
        const int g = tid / TP, t = tid - g * TP;
        const bool fft_thread = tid < FT;
        const int gl = fft_thread ? g : 0;
        const int kl = gl / WC, wcl = gl - kl * WC;
        const bool col_ok = fft_thread && qx0 + wcl < nx2;
        CT* line = s_buf + gl * LL;
        // branch-free: a thread without a column reads column qx0 and keeps zero
        const int64_t colofs = (int64_t)qy * nx2 + (col_ok ? qx0 + wcl : qx0);
        __syncthreads();  // twiddles
        for (int c = c0; c < c1; ++c) {
            const int64_t lofs = ((int64_t)c * K + kl) * nz * plane + colofs;
            if (fft_thread) {
                CT xr[NR];
#pragma unroll
                for (int p = 0; p < NR / R0; ++p)
#pragma unroll
                    for (int r = 0; r < R0; ++r) {
                        const int i = t + p * TP + r * NB0;
                        CT x = cz<CT>();
                        if (R0 % 2 != 0 || r < R0 / 2) {  // i < NZ2 / 2 = nz for even R0
                            const int ic = i < nz ? i : 0;
                            const CT y = st_load<CT, ST>(T2 + lofs + (int64_t)ic * plane);
                            x = (i < nz && col_ok) ? y : cz<CT>();
                        }
                        xr[p * R0 + r] = x;
                    }
                fft_wave<NZ2, -1, true, false>(line, s_tw, t, true, xr);
            }
            __syncthreads();
            multiply(WC * LL, [&](int wc) { return s_buf + wc * LL; });
            __syncthreads();
            if (fft_thread) {
                CT xr[NR];
                fft_wave<NZ2, +1, false, true>(line, s_tw, t, true, xr);
#pragma unroll
                for (int p = 0; p < NBL / TP; ++p)
#pragma unroll
                    for (int r = 0; r < RL; ++r) {
                        const int i = t + p * TP + r * NBL;
                        if ((RL % 2 != 0 || r < RL / 2) && i < nz && col_ok)
                            st_store<CT, ST>(T2 + lofs + (int64_t)i * plane, xr[p * RL + r]);
                    }
            }
        }
        return;
    }

    const int g = tid / TP, t = tid - g * TP;
    const bool fft_thread = tid < FT;
    for (int c = c0; c < c1; ++c) {
        // element (line = k*WC + wc, z): lanes run over wc fastest for coalescing
        for (int el = tid; el < WC * K * NZ2; el += BLK) {
            const int wc = el % WC;
            const int r = el / WC;
            const int k = r % K, z = r / K;
            const int qx = qx0 + wc;
            CT v = cz<CT>();
            if (z < nz && qx < nx2)
                v = st_load<CT, ST>(T2 + (((int64_t)c * K + k) * nz + z) * plane +
                                    (int64_t)qy * nx2 + qx);
            s_buf[(wc * K + k) * LL + lds_pad<NZ2, CT>(z)] = v;
        }
        __syncthreads();
        fft_lds<NZ2, -1>(s_buf + (fft_thread ? g : 0) * LL, s_tw, t, fft_thread);
#pragma unroll
        for (int e = 0; e < EV; ++e) {
            const int v = tid + e * BLK;
            if (v < NV) {
                const int wc = v / NZ2, qz = v - wc * NZ2;
                CT* col = s_buf + (size_t)wc * K * LL + lds_pad<NZ2, CT>(qz);
                CT in[K];
#pragma unroll
                for (int k = 0; k < K; ++k) in[k] = col[k * LL];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    CT acc = cz<CT>();
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        if constexpr (LREAL) {
                            const RT l = (RT)lr[e][j <= k ? packed_index(j, k) : packed_index(k, j)];
                            acc.x = fma(l, in[k].x, acc.x);
                            acc.y = fma(l, in[k].y, acc.y);
                        } else {
                            if (j <= k) cfma(acc, cto<CT>(lr[e][packed_index(j, k)]), in[k]);
                            else cfmac(acc, cto<CT>(lr[e][packed_index(k, j)]), in[k]);
                        }
                    }
                    col[j * LL] = acc;
                }
            }
        }
        __syncthreads();
        fft_lds<NZ2, +1>(s_buf + (fft_thread ? g : 0) * LL, s_tw, t, fft_thread);
        for (int el = tid; el < WC * K * NZ2; el += BLK) {
            const int wc = el % WC;
            const int r = el / WC;
            const int k = r % K, z = r / K;
            const int qx = qx0 + wc;
            if (z < nz && qx < nx2)
                st_store<CT, ST>(T2 + (((int64_t)c * K + k) * nz + z) * plane + (int64_t)qy * nx2 + qx,
                                 s_buf[(wc * K + k) * LL + lds_pad<NZ2, CT>(z)]);
        }
        __syncthreads();
    }
}

template <int NZ2, int K, bool LREAL, class CT, class ST>
void launch_pass_c_3d_t(const Apply3DArgs& a, hipStream_t st, const int* skip) {
    using S = PassC3Shape<NZ2, K, CT>;
    const int nx2 = 2 * a.nx, ny2 = 2 * a.ny;
    dim3 grid((nx2 + S::WC - 1) / S::WC, ny2, a.Gb);
    hipLaunchKernelGGL((pass_c_3d<NZ2, K, LREAL, CT, ST>), grid, dim3(S::BLK), 0, st,
                       static_cast<ST*>(a.T2), a.lam, static_cast<const CT*>(a.tw_z), nx2, ny2,
                       a.nz, a.C, a.Gb, skip);
}

template <int NZ2, int K, bool LREAL>
void launch_pass_c_3d_mode(const Apply3DArgs& a, hipStream_t st, const int* skip) {
    if (!a.f64) launch_pass_c_3d_t<NZ2, K, LREAL, float2, float2>(a, st, skip);
    else if (a.compact) launch_pass_c_3d_t<NZ2, K, LREAL, double2, cx64>(a, st, skip);
    else launch_pass_c_3d_t<NZ2, K, LREAL, double2, double2>(a, st, skip);
}

#define MSR_DECL_PASSC3(KK) void launch_pass_c_3d_k##KK(const Apply3DArgs& a, hipStream_t st, const int* skip);
MSR_DECL_PASSC3(1) MSR_DECL_PASSC3(2) MSR_DECL_PASSC3(3) MSR_DECL_PASSC3(4)
MSR_DECL_PASSC3(5) MSR_DECL_PASSC3(6) MSR_DECL_PASSC3(7) MSR_DECL_PASSC3(8)
#undef MSR_DECL_PASSC3

#define MSR_DEFINE_PASSC3(KK)                                                                  \
    void launch_pass_c_3d_k##KK(const Apply3DArgs& a, hipStream_t st, const int* skip) {       \
        const int nz2 = 2 * a.nz;                                                              \
        bool ok = dispatch_fft_len(nz2, [&]<int N>() {                                         \
            if constexpr (N <= 512) {                                                          \
                if (a.lam_real) launch_pass_c_3d_mode<N, KK, true>(a, st, skip);               \
                else launch_pass_c_3d_mode<N, KK, false>(a, st, skip);                         \
            } else {                                                                           \
                fail(MSR_ERR_UNSUPPORTED, "Toeplitz grid 2Nz = %d above 512", N);             \
            }                                                                                  \
        });                                                                                    \
        MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Nz = %d", nz2);           \
    }

}  // namespace msr
