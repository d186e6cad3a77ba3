// Kernels of the batched 1D FFT line passes (see fftlines.hpp), instantiated
// per length by the fftlines_p*.hip translation units.
#pragma once

#include "fft.hpp"
#include "fftlines.hpp"

#include <cstdlib>

namespace msr {

// LDS budget of a line-pass workgroup (gfx950: 160 KiB per CU), with room for
// the twiddle table
constexpr int LINE_LDS_BYTES = 160 * 1024;

__device__ __forceinline__ int select_bin(int j, int lout, int nf, int sel) {
    int r;
    if (sel == SEL_FIRST) return j;
    if (sel == SEL_CENTER) r = j - lout / 2;
    else r = j < (lout + 1) / 2 ? j : j - lout;
    return r < 0 ? r + nf : r;
}

// lines per workgroup: 8 (ComplexF64) or 16 (ComplexF32) up to 256 threads,
// halved until the lines and the twiddle table fit the LDS
template <int NF, class CT>
constexpr int rows_lines() {
    int f = sizeof(CT) == 16 ? 8 : 16;
    if (f * FFTPlan<NF>::T > 256) f = 256 / FFTPlan<NF>::T;
    while (f > 1 && (size_t)(f * (lds_line<NF, CT>() + 1) + NF) * sizeof(CT) > LINE_LDS_BYTES) f /= 2;
    return f;
}

template <int NF, class CT>
constexpr int rows_block() {
    return rows_lines<NF, CT>() * FFTPlan<NF>::T;
}

template <int NF, int DIR, class CT>
__global__ void __launch_bounds__(256) rows_kernel(LinePass p, const CT* __restrict__ tw) {
    if (p.skip && *p.skip) return;
    using RT = real_t<CT>;
    constexpr int TP = FFTPlan<NF>::T;
    constexpr int BLK = rows_block<NF, CT>();
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NF, CT>() + 1;
    // ComplexF64 twiddles from the global table (L1 / L2), ComplexF32 ones
    // from an LDS copy: the smaller LDS footprint admits more workgroups per CU
    __shared__ CT s_tw_lds[(sizeof(CT) == 16) ? 1 : NF];
    const CT* s_tw = (sizeof(CT) == 16) ? tw : s_tw_lds;
    __shared__ CT s[F * LD];
    const CT* in = static_cast<const CT*>(p.in);
    CT* out = static_cast<CT*>(p.out);
    if constexpr (!(sizeof(CT) == 16)) load_twiddles<NF>(s_tw_lds, tw);
    const int64_t l0 = (int64_t)blockIdx.x * F;
    // ComplexF32 lines in pairs of consecutive elements, one 16-byte access
    // each (even pitches and a 16-byte aligned base; a last odd element goes
    // alone), as the column passes pair adjacent columns
    constexpr bool PAIRED = sizeof(CT) == 8 && NF % 2 == 0;
    const bool pair_in = PAIRED && !p.single && (p.in_pitch & 1) == 0 &&
                         (reinterpret_cast<uintptr_t>(p.in) & 15u) == 0;
    const bool pair_out = PAIRED && !p.single && (p.out_pitch & 1) == 0 &&
                          (reinterpret_cast<uintptr_t>(p.out) & 15u) == 0;
    if (p.isel == SEL_FIRST && !p.win && !p.win64 && pair_in) {
        if constexpr (PAIRED) {
            for (int e = threadIdx.x; e < F * (NF / 2); e += BLK) {
                const int l = e / (NF / 2), i = 2 * (e - l * (NF / 2));
                const int64_t line = l0 + l;
                CT v0 = cz<CT>(), v1 = cz<CT>();
                if (line < p.nlines) {
                    const int64_t idx = line * p.in_pitch + i;
                    if (i + 1 < p.lin) {
                        const float4 w = *reinterpret_cast<const float4*>(in + idx);
                        v0 = make_float2(w.x, w.y);
                        v1 = make_float2(w.z, w.w);
                    } else if (i < p.lin) {
                        v0 = in[idx];
                    }
                }
                s[l * LD + lds_pad<NF, CT>(i)] = v0;
                s[l * LD + lds_pad<NF, CT>(i + 1)] = v1;
            }
        }
    } else if (p.isel == SEL_FIRST && !p.win && !p.win64) {
        for (int e = threadIdx.x; e < F * NF; e += BLK) {
            const int l = e / NF, i = e - l * NF;
            const int64_t line = l0 + l;
            CT v = cz<CT>();
            if (line < p.nlines && i < p.lin) v = in[line * p.in_pitch + i];
            s[l * LD + lds_pad<NF, CT>(i)] = v;
        }
    } else {
        for (int e = threadIdx.x; e < F * NF; e += BLK) s[(e / NF) * LD + lds_pad<NF, CT>(e % NF)] = cz<CT>();
        __syncthreads();
        for (int e = threadIdx.x; e < F * p.lin; e += BLK) {
            const int l = e / p.lin, i = e - l * p.lin;
            const int64_t line = l0 + l;
            if (line < p.nlines) {
                CT v = in[line * p.in_pitch + i];
                if (p.win64) v = cscale(v, (RT)p.win64[i]);
                else if (p.win) v = cscale(v, (RT)p.win[i]);
                s[l * LD + lds_pad<NF, CT>(select_bin(i, p.lin, NF, p.isel))] = v;
            }
        }
    }
    __syncthreads();
    const int g = threadIdx.x / TP, t = threadIdx.x - g * TP;
    fft_lds<NF, DIR>(s + g * LD, s_tw, t);
    if constexpr (PAIRED) {
        if (pair_out) {
            const int LP = (p.lout + 1) / 2;
            for (int e = threadIdx.x; e < F * LP; e += BLK) {
                const int l = e / LP, j = 2 * (e - l * LP);
                const int64_t line = l0 + l;
                if (line >= p.nlines) continue;
                CT v0 = s[l * LD + lds_pad<NF, CT>(select_bin(j, p.lout, NF, p.sel))];
                if (p.w64) v0 = cscale(v0, (RT)p.w64[j]);
                else if (p.w) v0 = cscale(v0, (RT)p.w[j]);
                const int64_t idx = line * p.out_pitch + j;
                if (j + 1 < p.lout) {
                    CT v1 = s[l * LD + lds_pad<NF, CT>(select_bin(j + 1, p.lout, NF, p.sel))];
                    if (p.w64) v1 = cscale(v1, (RT)p.w64[j + 1]);
                    else if (p.w) v1 = cscale(v1, (RT)p.w[j + 1]);
                    *reinterpret_cast<float4*>(out + idx) = make_float4(v0.x, v0.y, v1.x, v1.y);
                } else {
                    out[idx] = v0;
                }
            }
            return;
        }
    }
    for (int e = threadIdx.x; e < F * p.lout; e += BLK) {
        const int l = e / p.lout, j = e - l * p.lout;
        const int64_t line = l0 + l;
        if (line >= p.nlines) continue;
        CT v = s[l * LD + lds_pad<NF, CT>(select_bin(j, p.lout, NF, p.sel))];
        if (p.w64) v = cscale(v, (RT)p.w64[j]);
        else if (p.w) v = cscale(v, (RT)p.w[j]);
        out[line * p.out_pitch + j] = v;
    }
}

// columns per workgroup, halved until they fit the LDS
template <int NF, class CT>
constexpr int cols_wc() {
    int w = (NF <= 512 ? 16 : 8) / (sizeof(CT) == 16 ? 2 : 1);
    while (w > 1 && (size_t)(w * (lds_line<NF, CT>() + 1) + NF) * sizeof(CT) > LINE_LDS_BYTES) w /= 2;
    return w;
}

template <int NF, int DIR, class CT>
__global__ void __launch_bounds__(256) cols_kernel(LinePass p, const CT* __restrict__ tw) {
    if (p.skip && *p.skip) return;
    using RT = real_t<CT>;
    constexpr int TP = FFTPlan<NF>::T;
    constexpr int WC = cols_wc<NF, CT>();
    constexpr int BLK = (WC * TP) < 256 ? WC * TP : 256;
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NF, CT>() + 1;
    // ComplexF64 twiddles from the global table (L1 / L2), ComplexF32 ones
    // from an LDS copy: the smaller LDS footprint admits more workgroups per CU
    __shared__ CT s_tw_lds[(sizeof(CT) == 16) ? 1 : NF];
    const CT* s_tw = (sizeof(CT) == 16) ? tw : s_tw_lds;
    __shared__ CT s[WC * LD];
    const CT* in = static_cast<const CT*>(p.in);
    CT* out = static_cast<CT*>(p.out);
    // element idx of the input / output, ComplexF64 or (f64 only) cx64
    auto load_in = [&](int64_t idx) -> CT {
        if constexpr (sizeof(CT) == 16)
            if (p.in_compact) return st_load<CT, cx64>(static_cast<const cx64*>(p.in) + idx);
        return in[idx];
    };
    auto store_out = [&](int64_t idx, CT v) {
        if constexpr (sizeof(CT) == 16)
            if (p.out_compact) {
                st_store<CT, cx64>(static_cast<cx64*>(p.out) + idx, v);
                return;
            }
        out[idx] = v;
    };
    if constexpr (!(sizeof(CT) == 16)) load_twiddles<NF>(s_tw_lds, tw);
    // Workgroups go to the 8 XCDs round-robin by linear id, so neighbouring
    // column groups (the two halves of a 128-byte line when WC values are 64
    // bytes) would land on different L2s. Tile 2u+1 is therefore taken by the
    // workgroup 8 ids after tile 2u's, on the same XCD and dispatched next to
    // it (a bijection on each run of 16 ids; a ragged tail keeps its own).
    const unsigned nbx = gridDim.x, total = nbx * gridDim.y;
    unsigned L = blockIdx.x + nbx * blockIdx.y;
    if (L < total - total % 16u) L = (((L >> 4) << 3) + (L & 7u)) * 2u + ((L >> 3) & 1u);
    const int64_t col0 = (int64_t)(L % nbx) * WC;
    const int64_t b = L / nbx;
    // 8-byte values (compact ComplexF64, or ComplexF32) in pairs: one 16-byte
    // load or store covers two adjacent columns, half the memory instructions
    // of single values, with the same decode / encode per value.  Needs WC
    // even, even strides and a 16-byte aligned base, so that every pair (col0
    // is a multiple of WC) is aligned; a last odd column goes alone.
    constexpr bool PAIRED = WC % 2 == 0;
    bool pair_in = false, pair_out = false;
    if constexpr (PAIRED) {
        const bool in8 = sizeof(CT) == 8 || p.in_compact, out8 = sizeof(CT) == 8 || p.out_compact;
        pair_in = !p.single && in8 && ((p.in_stride | p.in_bstride) & 1) == 0 &&
                  (reinterpret_cast<uintptr_t>(p.in) & 15u) == 0;
        pair_out = !p.single && out8 && !p.lam && ((p.out_stride | p.out_bstride) & 1) == 0 &&
                   (reinterpret_cast<uintptr_t>(p.out) & 15u) == 0;
    }
    // the pair of columns at element idx, idx + 1 (pair_in / pair_out only)
    auto load_pair = [&](int64_t idx, CT& v0, CT& v1) {
        if constexpr (sizeof(CT) == 16) {
            const uint4 w = *reinterpret_cast<const uint4*>(static_cast<const cx64*>(p.in) + idx);
            v0 = cx64_decode(cx64{w.x, w.y});
            v1 = cx64_decode(cx64{w.z, w.w});
        } else {
            const float4 w = *reinterpret_cast<const float4*>(in + idx);
            v0 = make_float2(w.x, w.y);
            v1 = make_float2(w.z, w.w);
        }
    };
    auto store_pair = [&](int64_t idx, CT v0, CT v1) {
        if constexpr (sizeof(CT) == 16) {
            const cx64 c0 = cx64_encode(v0.x, v0.y), c1 = cx64_encode(v1.x, v1.y);
            *reinterpret_cast<uint4*>(static_cast<cx64*>(p.out) + idx) = make_uint4(c0.w0, c0.w1, c1.w0, c1.w1);
        } else {
            *reinterpret_cast<float4*>(out + idx) = make_float4(v0.x, v0.y, v1.x, v1.y);
        }
    };
    if (p.isel == SEL_FIRST && !p.win && !p.win64 && pair_in) {
        if constexpr (PAIRED) {
            for (int e = threadIdx.x; e < (WC / 2) * NF; e += BLK) {
                const int wc = 2 * (e % (WC / 2)), i = e / (WC / 2);
                const int64_t col = col0 + wc;
                CT v0 = cz<CT>(), v1 = cz<CT>();
                if (i < p.lin) {
                    const int64_t idx = b * p.in_bstride + (int64_t)i * p.in_stride + col;
                    if (col + 1 < p.ncols) load_pair(idx, v0, v1);
                    else if (col < p.ncols) v0 = load_in(idx);
                }
                s[wc * LD + lds_pad<NF, CT>(i)] = v0;
                s[(wc + 1) * LD + lds_pad<NF, CT>(i)] = v1;
            }
        }
    } else if (p.isel == SEL_FIRST && !p.win && !p.win64) {
        for (int e = threadIdx.x; e < WC * NF; e += BLK) {
            const int wc = e % WC, i = e / WC;
            const int64_t col = col0 + wc;
            CT v = cz<CT>();
            if (col < p.ncols && i < p.lin) v = load_in(b * p.in_bstride + i * p.in_stride + col);
            s[wc * LD + lds_pad<NF, CT>(i)] = v;
        }
    } else {
        for (int e = threadIdx.x; e < WC * NF; e += BLK) s[(e % WC) * LD + lds_pad<NF, CT>(e / WC)] = cz<CT>();
        __syncthreads();
        for (int e = threadIdx.x; e < WC * p.lin; e += BLK) {
            const int wc = e % WC, i = e / WC;
            const int64_t col = col0 + wc;
            if (col < p.ncols) {
                CT v = load_in(b * p.in_bstride + i * p.in_stride + col);
                if (p.win64) v = cscale(v, (RT)p.win64[i]);
                else if (p.win) v = cscale(v, (RT)p.win[i]);
                s[wc * LD + lds_pad<NF, CT>(select_bin(i, p.lin, NF, p.isel))] = v;
            }
        }
    }
    __syncthreads();
    const int g = threadIdx.x / TP, t = threadIdx.x - g * TP;
#pragma unroll
    for (int r0 = 0; r0 < WC; r0 += F) fft_lds<NF, DIR>(s + (r0 + g) * LD, s_tw, t);
    if constexpr (PAIRED) {
        if (pair_out) {
            for (int e = threadIdx.x; e < (WC / 2) * p.lout; e += BLK) {
                const int wc = 2 * (e % (WC / 2)), j = e / (WC / 2);
                const int64_t col = col0 + wc;
                if (col >= p.ncols) continue;
                const int bin = lds_pad<NF, CT>(select_bin(j, p.lout, NF, p.sel));
                CT v0 = s[wc * LD + bin], v1 = s[(wc + 1) * LD + bin];
                if (p.w64) {
                    v0 = cscale(v0, (RT)p.w64[j]);
                    v1 = cscale(v1, (RT)p.w64[j]);
                } else if (p.w) {
                    v0 = cscale(v0, (RT)p.w[j]);
                    v1 = cscale(v1, (RT)p.w[j]);
                }
                const int64_t idx = b * p.out_bstride + (int64_t)j * p.out_stride + col;
                if (col + 1 < p.ncols) store_pair(idx, v0, v1);
                else store_out(idx, v0);
            }
            return;
        }
    }
    for (int e = threadIdx.x; e < WC * p.lout; e += BLK) {
        const int wc = e % WC, j = e / WC;
        const int64_t col = col0 + wc;
        if (col >= p.ncols) continue;
        CT v = s[wc * LD + lds_pad<NF, CT>(select_bin(j, p.lout, NF, p.sel))];
        if (p.w64) v = cscale(v, (RT)p.w64[j]);
        else if (p.w) v = cscale(v, (RT)p.w[j]);
        if (!p.lam) {
            store_out(b * p.out_bstride + j * p.out_stride + col, v);
        } else {
            // dense Toeplitz kernel, rounded once to ComplexF32 / Float32
            const int64_t pp = b + p.p0;
            int64_t off, mi;
            if (!p.lam3d) {
                off = (col * p.P + pp) * p.lout + j;
                mi = (int64_t)j * p.nx2 + col;
            } else {
                off = (col * p.P + pp) * p.lout + j;           // col = qy*nx2 + qx
                mi = (int64_t)j * p.ny2 * p.nx2 + col;         // j = qz
            }
            const bool on = p.mask[mi] != 0;
            if (p.lam_real) reinterpret_cast<float*>(p.lam)[off] = on ? (float)v.x : 0.f;
            else reinterpret_cast<float2*>(p.lam)[off] = on ? make_float2((float)v.x, (float)v.y)
                                                           : make_float2(0.f, 0.f);
        }
    }
}

// MSR_COLS_SINGLE=1 in the environment: 8-byte line and column values one
// per access (to compare with the paired accesses)
inline bool cols_single_forced() {
    const char* e = std::getenv("MSR_COLS_SINGLE");
    return e != nullptr && e[0] == '1';
}

template <int N, class CT>
void launch_rows_n(const LinePass& p, const CT* tw, hipStream_t st) {
    constexpr int BLK = rows_block<N, CT>();
    constexpr int F = BLK / FFTPlan<N>::T;
    dim3 grid((unsigned)ceil_div(p.nlines, F));
    LinePass q = p;
    q.single = cols_single_forced() ? 1 : 0;
    if (q.dir < 0) hipLaunchKernelGGL((rows_kernel<N, -1, CT>), grid, dim3(BLK), 0, st, q, tw);
    else hipLaunchKernelGGL((rows_kernel<N, 1, CT>), grid, dim3(BLK), 0, st, q, tw);
}

template <int N, class CT>
void launch_cols_n(const LinePass& p0, const CT* tw, hipStream_t st) {
    LinePass p = p0;
    p.single = cols_single_forced() ? 1 : 0;
    constexpr int WC = cols_wc<N, CT>();
    constexpr int BLK = (WC * FFTPlan<N>::T) < 256 ? WC * FFTPlan<N>::T : 256;
    dim3 grid((unsigned)ceil_div(p.ncols, WC), (unsigned)p.nb);
    if (p.dir < 0) hipLaunchKernelGGL((cols_kernel<N, -1, CT>), grid, dim3(BLK), 0, st, p, tw);
    else hipLaunchKernelGGL((cols_kernel<N, 1, CT>), grid, dim3(BLK), 0, st, p, tw);
}

// Launch the rows (or cols) pass of p if its length is one of LIST: the body
// of line_pass_part<I> in each fftlines_p*.hip.
#define MSR_LINE_PART_BODY(LIST)                                                      \
    switch (p.nf) {                                                                   \
        LIST(MSR_LINE_CASE)                                                           \
        default: return false;                                                        \
    }
#define MSR_LINE_CASE(NN)                                                             \
    case NN:                                                                          \
        if (p.f64) {                                                                  \
            if (rows) launch_rows_n<NN, double2>(p, static_cast<const double2*>(tw), st); \
            else launch_cols_n<NN, double2>(p, static_cast<const double2*>(tw), st);  \
        } else {                                                                      \
            if (rows) launch_rows_n<NN, float2>(p, static_cast<const float2*>(tw), st);   \
            else launch_cols_n<NN, float2>(p, static_cast<const float2*>(tw), st);    \
        }                                                                             \
        return true;

}  // namespace msr
