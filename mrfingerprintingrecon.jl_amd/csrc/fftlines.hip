// Generic batched 1D FFT line passes (see fftlines.hpp).
#include "fft.hpp"
#include "fftlines.hpp"

namespace msr {

__device__ __forceinline__ int select_bin(int j, int lout, int nf, int sel) {
    int r;
    if (sel == SEL_FIRST) return j;
    if (sel == SEL_CENTER) r = j - lout / 2;
    else r = j < (lout + 1) / 2 ? j : j - lout;
    return r < 0 ? r + nf : r;
}

template <int NF>
constexpr int rows_block() {
    return (16 * FFTPlan<NF>::T) < 256 ? 16 * FFTPlan<NF>::T : 256;
}

template <int NF, int DIR>
__global__ void __launch_bounds__(256) rows_kernel(LinePass p, const float2* __restrict__ tw) {
    if (p.skip && *p.skip) return;
    constexpr int TP = FFTPlan<NF>::T;
    constexpr int BLK = rows_block<NF>();
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NF>() + 1;
    __shared__ float2 s_tw[NF];
    __shared__ float2 s[F * LD];
    load_twiddles<NF>(s_tw, tw);
    const int64_t l0 = (int64_t)blockIdx.x * F;
    if (p.isel == SEL_FIRST && !p.win) {
        for (int e = threadIdx.x; e < F * NF; e += BLK) {
            const int l = e / NF, i = e - l * NF;
            const int64_t line = l0 + l;
            float2 v = make_float2(0.f, 0.f);
            if (line < p.nlines && i < p.lin) v = p.in[line * p.in_pitch + i];
            s[l * LD + lds_pad(i)] = v;
        }
    } else {
        for (int e = threadIdx.x; e < F * NF; e += BLK) s[(e / NF) * LD + lds_pad(e % NF)] = make_float2(0.f, 0.f);
        __syncthreads();
        for (int e = threadIdx.x; e < F * p.lin; e += BLK) {
            const int l = e / p.lin, i = e - l * p.lin;
            const int64_t line = l0 + l;
            if (line < p.nlines) {
                float2 v = p.in[line * p.in_pitch + i];
                if (p.win) v = cscale(v, p.win[i]);
                s[l * LD + lds_pad(select_bin(i, p.lin, NF, p.isel))] = v;
            }
        }
    }
    __syncthreads();
    const int g = threadIdx.x / TP, t = threadIdx.x - g * TP;
    fft_lds<NF, DIR>(s + g * LD, s_tw, t);
    for (int e = threadIdx.x; e < F * p.lout; e += BLK) {
        const int l = e / p.lout, j = e - l * p.lout;
        const int64_t line = l0 + l;
        if (line >= p.nlines) continue;
        float2 v = s[l * LD + lds_pad(select_bin(j, p.lout, NF, p.sel))];
        if (p.w) v = cscale(v, p.w[j]);
        p.out[line * p.out_pitch + j] = v;
    }
}

template <int NF>
constexpr int cols_wc() {
    return NF <= 512 ? 16 : 8;
}

template <int NF, int DIR>
__global__ void __launch_bounds__(256) cols_kernel(LinePass p, const float2* __restrict__ tw) {
    if (p.skip && *p.skip) return;
    constexpr int TP = FFTPlan<NF>::T;
    constexpr int WC = cols_wc<NF>();
    constexpr int BLK = (WC * TP) < 256 ? WC * TP : 256;
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NF>() + 1;
    __shared__ float2 s_tw[NF];
    __shared__ float2 s[WC * LD];
    load_twiddles<NF>(s_tw, tw);
    const int64_t col0 = (int64_t)blockIdx.x * WC;
    const int64_t b = blockIdx.y;
    if (p.isel == SEL_FIRST && !p.win) {
        for (int e = threadIdx.x; e < WC * NF; e += BLK) {
            const int wc = e % WC, i = e / WC;
            const int64_t col = col0 + wc;
            float2 v = make_float2(0.f, 0.f);
            if (col < p.ncols && i < p.lin) v = p.in[b * p.in_bstride + i * p.in_stride + col];
            s[wc * LD + lds_pad(i)] = v;
        }
    } else {
        for (int e = threadIdx.x; e < WC * NF; e += BLK) s[(e % WC) * LD + lds_pad(e / WC)] = make_float2(0.f, 0.f);
        __syncthreads();
        for (int e = threadIdx.x; e < WC * p.lin; e += BLK) {
            const int wc = e % WC, i = e / WC;
            const int64_t col = col0 + wc;
            if (col < p.ncols) {
                float2 v = p.in[b * p.in_bstride + i * p.in_stride + col];
                if (p.win) v = cscale(v, p.win[i]);
                s[wc * LD + lds_pad(select_bin(i, p.lin, NF, p.isel))] = v;
            }
        }
    }
    __syncthreads();
    const int g = threadIdx.x / TP, t = threadIdx.x - g * TP;
#pragma unroll
    for (int r0 = 0; r0 < WC; r0 += F) fft_lds<NF, DIR>(s + (r0 + g) * LD, s_tw, t);
    for (int e = threadIdx.x; e < WC * p.lout; e += BLK) {
        const int wc = e % WC, j = e / WC;
        const int64_t col = col0 + wc;
        if (col >= p.ncols) continue;
        float2 v = s[wc * LD + lds_pad(select_bin(j, p.lout, NF, p.sel))];
        if (p.w) v = cscale(v, p.w[j]);
        if (!p.lam) {
            p.out[b * p.out_bstride + j * p.out_stride + col] = v;
        } else {
            int64_t off, mi;
            if (!p.lam3d) {
                off = (col * p.P + b) * p.lout + j;
                mi = (int64_t)j * p.nx2 + col;
            } else {
                off = (col * p.P + b) * p.lout + j;            // col = qy*nx2 + qx
                mi = (int64_t)j * p.ny2 * p.nx2 + col;         // j = qz
            }
            const bool on = p.mask[mi] != 0;
            if (p.lam_real) reinterpret_cast<float*>(p.lam)[off] = on ? v.x : 0.f;
            else reinterpret_cast<float2*>(p.lam)[off] = on ? v : make_float2(0.f, 0.f);
        }
    }
}

void run_rows(const LinePass& p, msr_ctx* ctx, hipStream_t st) {
    const float2* tw = twiddle_table(ctx, p.nf);
    bool ok = dispatch_fft_len(p.nf, [&]<int N>() {
        constexpr int BLK = rows_block<N>();
        constexpr int F = BLK / FFTPlan<N>::T;
        dim3 grid((unsigned)ceil_div(p.nlines, F));
        if (p.dir < 0) hipLaunchKernelGGL((rows_kernel<N, -1>), grid, dim3(BLK), 0, st, p, tw);
        else hipLaunchKernelGGL((rows_kernel<N, 1>), grid, dim3(BLK), 0, st, p, tw);
    });
    MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "no FFT kernel for length %d", p.nf);
    MSR_CHECK_LAUNCH();
}

void run_cols(const LinePass& p, msr_ctx* ctx, hipStream_t st) {
    const float2* tw = twiddle_table(ctx, p.nf);
    bool ok = dispatch_fft_len(p.nf, [&]<int N>() {
        constexpr int WC = cols_wc<N>();
        constexpr int BLK = (WC * FFTPlan<N>::T) < 256 ? WC * FFTPlan<N>::T : 256;
        dim3 grid((unsigned)ceil_div(p.ncols, WC), (unsigned)p.nb);
        if (p.dir < 0) hipLaunchKernelGGL((cols_kernel<N, -1>), grid, dim3(BLK), 0, st, p, tw);
        else hipLaunchKernelGGL((cols_kernel<N, 1>), grid, dim3(BLK), 0, st, p, tw);
    });
    MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "no FFT kernel for length %d", p.nf);
    MSR_CHECK_LAUNCH();
}

}  // namespace msr
