// Device pieces shared by the 2D and 3D output-owned backprojection gathers
// (spread_bp.hip, spread_bp3.hip): the spreading kernel, the per-point
// channel values and the DPP-broadcast multiply-adds.
#pragma once

#include "spread.hpp"

namespace msr {

constexpr int ITEM_PAIRS = 1024;   // (block, point) pairs per work item
constexpr size_t BP_VALS_BUDGET = (size_t)1 << 30;  // bytes of channel values per pass
constexpr int BP_CGMAX = 16;       // largest channel group (2 CG value floats in two VGPRs)
constexpr int BP_DEPTH = 4;        // points whose loads are in flight per wavefront

__device__ __forceinline__ float es_phi_f(float t) {
    const float z = t * (2.0f / SPREAD_W);
    const float a = 1.0f - z * z;
    return a > 0.0f ? expf(SPREAD_BETA * (sqrtf(a) - 1.0f)) : 0.0f;
}

__device__ __forceinline__ float2 basis_val(const ValueSrc& s, int t, int k) {
    if (!s.U) return make_float2(t == k ? 1.f : 0.f, 0.f);
    if (s.u_real) return make_float2(reinterpret_cast<const float*>(s.U)[t + (int64_t)k * s.nt], 0.f);
    return reinterpret_cast<const float2*>(s.U)[t + (int64_t)k * s.nt];
}

// vals[gl][s][j] = conj(U[t_s, k]) y[jsel(s), c] dcf(s) for channel
// ch = (g0 + gl) CG + j = c K + k (zero beyond the last channel): the channel
// values of sorted point s, CG per group; one thread per (point, group),
// blockIdx.y = gl, each coil sample read once per group.
template <int CG>
__global__ void bp_values_kernel(ValueSrc src, const int64_t* __restrict__ jsel_s,
                                 const float2* __restrict__ kc_s, const float* __restrict__ r2_s,
                                 const int* __restrict__ perm,
                                 const int* __restrict__ frame_s, int64_t M, int nch, int g0,
                                 float2* __restrict__ vals) {
    const int gl = blockIdx.y;
    const int chb = (g0 + gl) * CG;
    const int K = src.K;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < M;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t js = jsel_s[s];
        const int t = frame_s[s];
        float dcf = 1.f;
        if (src.dcf_mode == 1 || src.dcf_mode == 2) {
            // |k|^2 of the sample: from its 2D coordinates, or given (3D)
            float r2;
            if (r2_s) {
                r2 = r2_s[s];
            } else {
                const float2 q = kc_s[s];
                r2 = q.x * q.x + q.y * q.y;
            }
            dcf = src.dcf_mode == 1 ? sqrtf(r2) : r2;
        } else if (src.dcf_mode == 3) {
            dcf = src.dcf[perm[s]];
        }
        float2 out[CG];
        int c = chb / K, k = chb - c * K;
        float2 y = make_float2(0.f, 0.f);
        int cy = -1;
#pragma unroll
        for (int j = 0; j < CG; ++j) {
            out[j] = make_float2(0.f, 0.f);
            if (chb + j < nch) {
                if (c != cy) {
                    y = cscale(src.data[js + (int64_t)c * src.nsel], dcf);
                    cy = c;
                }
                out[j] = cmulc(basis_val(src, t, k), y);
            }
            if (++k == K) {
                k = 0;
                ++c;
            }
        }
        if constexpr (CG % 2 == 0) {
            float4* dst = reinterpret_cast<float4*>(vals + ((int64_t)gl * M + s) * CG);
#pragma unroll
            for (int j = 0; j < CG / 2; ++j)
                dst[j] = make_float4(out[2 * j].x, out[2 * j].y, out[2 * j + 1].x, out[2 * j + 1].y);
        } else {
#pragma unroll
            for (int j = 0; j < CG; ++j) vals[((int64_t)gl * M + s) * CG + j] = out[j];
        }
    }
}

// acc[J0 + j] += v_j w for j < 16, where v_j is lane j of the 16-lane row of v
// (the point's values, the same 16 floats in each of the four rows): DPP
// row_newbcast hands lane j of each row to the whole row, so the 16 FMAs need
// no other instruction.  s_nop 1: two wait states before a DPP read of a VGPR
// that a VALU instruction may have just written.
template <int J0, int NA>
__device__ __forceinline__ void fma_bcast16(float (&acc)[NA], float v, float w) {
    asm volatile(
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %16, %17 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %16, %17 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %2, %16, %17 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %3, %16, %17 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %4, %16, %17 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %5, %16, %17 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %6, %16, %17 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %7, %16, %17 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %8, %16, %17 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %9, %16, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %10, %16, %17 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %11, %16, %17 row_newbcast:11 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %12, %16, %17 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %13, %16, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %14, %16, %17 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %15, %16, %17 row_newbcast:15 row_mask:0xf bank_mask:0xf"
        : "+v"(acc[J0 + 0]), "+v"(acc[J0 + 1]), "+v"(acc[J0 + 2]), "+v"(acc[J0 + 3]),
          "+v"(acc[J0 + 4]), "+v"(acc[J0 + 5]), "+v"(acc[J0 + 6]), "+v"(acc[J0 + 7]),
          "+v"(acc[J0 + 8]), "+v"(acc[J0 + 9]), "+v"(acc[J0 + 10]), "+v"(acc[J0 + 11]),
          "+v"(acc[J0 + 12]), "+v"(acc[J0 + 13]), "+v"(acc[J0 + 14]), "+v"(acc[J0 + 15])
        : "v"(v), "v"(w));
}

// The same for a single channel: acc[0..1] += (lane 0, lane 1 of the row) w.
__device__ __forceinline__ void fma_bcast2(float (&acc)[2], float v, float w) {
    asm volatile(
        "s_nop 1\n\t"
        "v_fmac_f32_dpp %0, %2, %3 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f32_dpp %1, %2, %3 row_newbcast:1 row_mask:0xf bank_mask:0xf"
        : "+v"(acc[0]), "+v"(acc[1])
        : "v"(v), "v"(w));
}

// channel group size: the fewest groups of at most BP_CGMAX channels, each
// rounded up to an instantiated size (1, 8 or 16)
inline int bp_group_size(int nch) {
    const int ng = (int)ceil_div(nch, BP_CGMAX);
    const int need = (int)ceil_div(nch, ng);
    return need == 1 ? 1 : (need <= 8 ? 8 : BP_CGMAX);
}

// channel groups per pass: all of them (one launch with enough wavefronts to
// fill the GPU), unless their values exceed BP_VALS_BUDGET
inline int bp_groups_per_pass(int64_t M, int CG, int ng) {
    const size_t per = sizeof(float2) * (size_t)std::max<int64_t>(M, 1) * CG;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)ng, BP_VALS_BUDGET / per));
}

}  // namespace msr
