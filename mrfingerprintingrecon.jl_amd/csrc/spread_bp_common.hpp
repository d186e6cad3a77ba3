// Device pieces shared by the 2D and 3D output-owned backprojection gathers
// (spread_bp.hip, spread_bp3.hip): the spreading kernel, the per-point
// channel values, the DPP-broadcast multiply-adds and the matrix-core gather.
#pragma once

#include <algorithm>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "spread.hpp"

namespace msr {

constexpr int ITEM_PAIRS = 1024;   // (block, point) pairs per work item
constexpr size_t BP_VALS_BUDGET = (size_t)1 << 30;  // bytes of channel values per pass
constexpr int BP_CGMAX = 16;       // largest channel group (32 value floats: one MFMA column block)
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
// values of sorted point s, CG per group; blockIdx.y = gl.  For even CG, CG / 2
// threads per point, each forming two channels and storing them as one 16-byte
// word, so a wavefront's stores cover consecutive bytes; for CG = 1 one thread
// per point.
constexpr int bp_values_lanes(int CG) { return CG % 2 == 0 ? CG / 2 : 1; }

template <int CG>
__global__ void bp_values_kernel(ValueSrc src, const int64_t* __restrict__ jsel_s,
                                 const float2* __restrict__ kc_s, const float* __restrict__ r2_s,
                                 const int* __restrict__ perm,
                                 const int* __restrict__ frame_s, int64_t M, int nch, int g0,
                                 float2* __restrict__ vals) {
    constexpr int LP = bp_values_lanes(CG);   // threads per point
    constexpr int CL = CG / LP;               // channels per thread
    const int gl = blockIdx.y;
    const int K = src.K;
    const int64_t total = M * LP;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = e / LP;
        const int j0 = (int)(e - s * LP) * CL;
        const int chb = (g0 + gl) * CG + j0;
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
        float2 out[CL];
        int c = chb / K, k = chb - c * K;
        float2 y = make_float2(0.f, 0.f);
        int cy = -1;
#pragma unroll
        for (int j = 0; j < CL; ++j) {
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
        float2* dst = vals + ((int64_t)gl * M + s) * CG + j0;
        if constexpr (CL == 2)
            *reinterpret_cast<float4*>(dst) = make_float4(out[0].x, out[0].y, out[1].x, out[1].y);
        else
            dst[0] = out[0];
    }
}

// grid of bp_values_kernel for M points (blockIdx.y: the groups)
inline int bp_values_grid(int64_t M, int CG) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(M * bp_values_lanes(CG), 256), 32768));
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

// ---------------------------------------------------------------- matrix cores
// A 16-channel group of an output block of 64 cells is the product
// S = V W over the block's points: V (32 x P) holds each point's 32 value
// floats (16 complex channels), W (P x 64) its kernel weight at each cell.
// v_mfma_f32_32x32x2f32 multiplies 32 x 2 by 2 x 32: lane l supplies row
// (l & 31) of A for point (l >> 5) and column (l & 31) of B for the same point,
// and holds S[i][l & 31] in register r for i = 8 (r / 4) + 4 (l >> 5) + r % 4.
// So a wavefront takes two points per step, with two instructions: one for
// cells 0..31 (acc0) and one for cells 32..63 (acc1), built in registers from
// the pair's per-dimension weights (Cells::load and Cells::product).  Compared
// with the broadcast FMAs above (2 CG FMAs per cell and point on the vector
// ALU), the matrix cores do the same multiply-adds at twice the rate of
// unpacked FP32 FMAs, and each lane loads one value float per step instead of
// two value VGPRs per point.
using f32x16 = float __attribute__((ext_vector_type(16)));

// s_waitcnt immediate (gfx9 encoding) that waits for every vector memory load:
// vmcnt 0, expcnt 7 and lgkmcnt 15 (no wait on those)
constexpr int BP_VMCNT0 = 0x0F70;

// channel of register pair (r, r + 1) (real, imaginary part) in lane half h
__host__ __device__ constexpr int mfma_channel(int r, int h) {
    return 4 * (r >> 2) + 2 * h + ((r & 3) >> 1);
}

// pair records past the last pair (zero weights, pair word 0) that the
// gathers' prefetch may read
constexpr int BP_PAD_PAIRS = 8 * BP_DEPTH;
// groups of at most this many points address their values with 32-bit offsets
constexpr int64_t BP_NARROW_POINTS = (int64_t)1 << 25;

// the 64-bit value offsets: for groups above BP_NARROW_POINTS points, and on
// request (environment MSR_BP_WIDE_OFFSETS=1) to test them at small sizes
inline bool bp_wide_offsets(int64_t M) {
    if (M > BP_NARROW_POINTS) return true;
    const char* e = std::getenv("MSR_BP_WIDE_OFFSETS");
    return e != nullptr && e[0] == '1';
}

__device__ __forceinline__ float bp_ldf(const char* __restrict__ base, uint32_t off) {
    return *reinterpret_cast<const float*>(base + (size_t)off);
}
__device__ __forceinline__ int bp_ldi(const char* __restrict__ base, uint32_t off) {
    return *reinterpret_cast<const int*>(base + (size_t)off);
}

// Walks the pairs q0 .. q1 - 1 (q1 > q0) of a block, two per step (lane
// half h = lane >> 5 takes pair q0 + 2 s + h), the loads of a step issued
// BP_DEPTH steps ahead of its use and the pair word BP_DEPTH steps before that
// (vector loads complete in order).  vals points at value float 0 of point 0
// of the group (32 floats per point).  All addresses are a uniform base plus a
// 32-bit lane offset (plus a constant), so a load costs no address arithmetic
// beyond one add per iteration and stage group: the pair words and weights
// are addressed from the item's first pair (the prefetch runs up to
// BP_PAD_PAIRS past the item's last pair, into the next items or the zero
// padding, and never uses what it reads there), the values from the group's
// point 0 (WIDE: 64-bit offsets, for groups above 4 GiB).  In the last step
// of an odd count the upper half's value is zero.  The sum of each cell runs
// over the pairs in a fixed order.
template <bool WIDE, class Cells>
__device__ __forceinline__ void mfma_gather16(const float* __restrict__ vals,
                                              const int* __restrict__ pairs,
                                              const float* __restrict__ pw, int q0, int q1,
                                              const Cells& cells, f32x16& acc0, f32x16& acc1) {
    constexpr int D = BP_DEPTH;
    constexpr uint32_t PSTEP = 2 * Cells::PW_BYTES;   // weight bytes per step
    const int lane = (int)threadIdx.x;
    const int h = lane >> 5;
    const int n = q1 - q0;
    __builtin_assume(n >= 1);
    const int steps = (n + 1) >> 1;
    const char* __restrict__ pwi =
        reinterpret_cast<const char*>(pw) + (size_t)q0 * Cells::PW_BYTES;
    const char* __restrict__ pqi = reinterpret_cast<const char*>(pairs + q0);
    const char* __restrict__ vb = reinterpret_cast<const char*>(vals);
    const uint32_t vlane = (uint32_t)(lane & 31) * 4;
    int wd[D];                    // pair words of steps s + D .. s + 2D - 1
    float av[D];                  // value floats of steps s .. s + D - 1
    float w[D][Cells::NW];        // weights of steps s .. s + D - 1
    auto value = [&](int word) -> float {
        if constexpr (WIDE)
            return vals[(size_t)(uint32_t)word * 32 + (lane & 31)];
        else
            return bp_ldf(vb, ((uint32_t)word << 7) + vlane);
    };
    // lane offsets: pair record of step 0, pair word of step 0
    const uint32_t orec0 = (uint32_t)h * Cells::PW_BYTES;
    const uint32_t oq0 = (uint32_t)h * 4;
#pragma unroll
    for (int i = 0; i < D; ++i) wd[i] = bp_ldi(pqi + 8 * i, oq0);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        av[i] = value(wd[i]);
        cells.load(pwi + i * PSTEP, orec0, w[i]);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) wd[i] = bp_ldi(pqi + 8 * (D + i), oq0);
    // The value loads of the first steps wait for their pair words, so they
    // leave the prologue in another order than the loop issues its loads; with
    // the prologue's loads complete, the waits at the loop head follow the
    // loop's own order (loads of the next D - 1 steps stay in flight).
    asm volatile("" ::: "memory");  // the prologue's loads stay above the wait
    __builtin_amdgcn_s_waitcnt(BP_VMCNT0);
    uint32_t orec = orec0;        // lane's record offset of step s (from pwi + D PSTEP)
    uint32_t oq = oq0;            // lane's word offset of step s (from pqi + 16 D)
    for (int s = 0; s < steps; s += D) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (s + i < steps) {
                float b0, b1;
                cells.product(w[i], b0, b1);
                float v = av[i];
                if ((n & 1) && s + i == steps - 1) v = h ? 0.f : v;
                acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(v, b0, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(v, b1, acc1, 0, 0, 0);
            }
            av[i] = value(wd[i]);
            cells.load(pwi + (D + i) * PSTEP, orec, w[i]);
            // the next pair word of stage i goes into the register of the
            // one just used (loaded any earlier, it takes another register and
            // a copy at the loop's end that waits for every load in flight)
            __builtin_amdgcn_sched_barrier(0);
            wd[i] = bp_ldi(pqi + 8 * (2 * D + i), oq);
        }
        orec += D * PSTEP;
        oq += 8 * D;
    }
}

// ------------------------------------------------------------ dispatch order
// A launch's workgroup i runs on XCD i mod BP_XCDS, and each XCD has its own
// L2.  In block order the neighbouring blocks that read the same points'
// values land on different XCDs, and the channel groups of a block (which
// read the same pair words and weights) run a whole pass apart.  The gathers
// therefore take their work from a dispatch table: the work items in Morton
// order of their blocks (the long items first), cut into BP_XCDS * BP_SEGMENTS
// segments of about equal work (pairs + BP_ITEM_COST per item), segment s on
// XCD s mod BP_XCDS, the channel groups of an item dispatched one after the
// other.  Each XCD then works through compact regions of blocks.  The result does not depend
// on the order (each item writes its own cells or partial block).
constexpr int BP_XCDS = 8;
constexpr int BP_SEGMENTS = 32;
constexpr int BP_ITEM_COST = 16;

// Morton key (bit interleave) of d block coordinates
inline uint64_t bp_morton_key(const int* c, int d) {
    uint64_t k = 0;
    for (int b = 0; b < 21; ++b)
        for (int a = 0; a < d; ++a) k |= (uint64_t)((c[a] >> b) & 1) << (b * d + a);
    return k;
}

// order[x * len + p]: the item at dispatch position p on XCD x, -1 past the
// XCD's last item; key[i]: the Morton key of item i's block
inline std::vector<int> bp_dispatch_order(const std::vector<int4>& items,
                                          const std::vector<uint64_t>& key, int64_t& len) {
    const int64_t n = (int64_t)items.size();
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    // items of at least ITEM_PAIRS / 2 pairs (the dense centre of a radial
    // trajectory) first, so that no long wavefront starts near the launch's end
    auto heavy = [&](int a) { return items[a].z - items[a].y >= ITEM_PAIRS / 2; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        if (heavy(a) != heavy(b)) return heavy(a);
        return key[a] < key[b];
    });
    double total = 0.0;
    for (const int4& it : items) total += it.z - it.y + BP_ITEM_COST;
    const int64_t nseg = std::max<int64_t>(1, std::min<int64_t>(BP_XCDS * BP_SEGMENTS, n));
    std::vector<std::vector<int>> lists(BP_XCDS);
    double acc = 0.0;
    int64_t seg = 0;
    for (const int i : idx) {
        lists[seg % BP_XCDS].push_back(i);
        acc += items[i].z - items[i].y + BP_ITEM_COST;
        while (seg + 1 < nseg && acc >= total * (double)(seg + 1) / (double)nseg) ++seg;
    }
    len = 0;
    for (const auto& l : lists) len = std::max<int64_t>(len, (int64_t)l.size());
    std::vector<int> order((size_t)(BP_XCDS * std::max<int64_t>(len, 1)), -1);
    for (int x = 0; x < BP_XCDS; ++x)
        for (size_t p = 0; p < lists[x].size(); ++p) order[(size_t)x * len + p] = lists[x][p];
    return order;
}

// the item (-1: a padding workgroup) and channel group (gl) of this workgroup
// in a launch of BP_XCDS * len * ngp workgroups
__device__ __forceinline__ int bp_item_of(const int* __restrict__ order, int64_t len, int ngp,
                                          int& gl) {
    const int64_t L = blockIdx.x;
    const int64_t x = L % BP_XCDS, seq = L / BP_XCDS;
    const int64_t p = seq / ngp;
    gl = (int)(seq - p * ngp);
    return order[x * len + p];
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
