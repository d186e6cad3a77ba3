// 3D ComplexF32 spreading of the backprojection (C*K value channels; the
// reference runs one type-1 NUFFT per coil and coefficient,
// ext/MRISubspaceReconCUDAExt/BackProjection.jl:55-71): the 2D gather of
// spread_bp.hip on 4 x 4 x 4 output blocks.
//
// One wavefront per (work item, group of CG channels) walks the item's points
// and sums wx wy wz v into the block's 64 cells: on the matrix cores for
// groups of 16 channels (mfma_gather16), else one lane per cell (lx, ly, lz)
// holding its cell's CG channel sums, with the point's values broadcast from
// lanes by DPP (two CG multiply-adds per cell and point).  The kernel weights of a point over the block's four
// columns, rows and planes (12 floats per (block, point) pair) are laid out
// when the streams are built.  A point's W = 8 support covers two or three
// blocks per dimension, about 21 pairs per point on average, and 512 of the
// ~1300 lane-cells its pairs visit.  Each cell sums its points in the sorted
// point order (bit-reproducible); blocks with more than ITEM_PAIRS pairs (the
// dense centre of a koosh ball) are split into items whose partial blocks a
// second kernel sums in item order.
//
// Built once per trajectory and grid (on the plan): each sorted point's cell,
// frame, data index and |k|^2; the pair lists (stable radix sort of (block,
// point) pairs by block); the pair weights.  Per call: the channel values of
// every sorted point (bp_values_kernel), then the gather.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "spread_bp_common.hpp"

namespace msr {

constexpr int BP3_TB = 4;  // block edge (cells): 4^3 = 64 lanes

void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

struct StreamGeom3 {
    int64_t n[3];
    int nb[3];
};

// distinct blocks holding the cells c-3 .. c+4 (periodic) of one dimension:
// two or three, four when the support wraps across a partial last block
// (n not a multiple of 4); t holds up to SUPPORT_MAX
constexpr int SUPPORT_MAX = 8;
__device__ __forceinline__ int support_blocks4(int64_t c, int64_t n, int* t) {
    int cnt = 0;
    for (int o = -3; o <= 4; ++o) {
        int64_t a = c + o;
        if (a < 0) a += n;
        if (a >= n) a -= n;
        const int tt = (int)(a / BP3_TB);
        bool seen = false;
        for (int i = 0; i < cnt; ++i) seen |= (t[i] == tt);
        if (!seen) t[cnt++] = tt;
    }
    return cnt;
}

__global__ void bp3_points_kernel(const float* __restrict__ kc, const int* __restrict__ frame,
                                  const int64_t* __restrict__ sel, const int* __restrict__ perm,
                                  int64_t M, StreamGeom3 g, float4* __restrict__ frac,
                                  int4* __restrict__ cell, int* __restrict__ frame_s,
                                  int64_t* __restrict__ jsel_s, float* __restrict__ r2_s,
                                  int* __restrict__ npairs) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < M;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[s];
        int64_t c[3];
        float f[3];
        int cnt = 1;
        for (int a = 0; a < 3; ++a) {
            const double x = (double)kc[j * 3 + a] * (double)g.n[a];
            const double fl = floor(x);
            f[a] = (float)(x - fl);
            int64_t cc = (int64_t)fl % g.n[a];
            if (cc < 0) cc += g.n[a];
            c[a] = cc;
            int t[SUPPORT_MAX];
            cnt *= support_blocks4(cc, g.n[a], t);
        }
        frac[s] = make_float4(f[0], f[1], f[2], 0.f);
        cell[s] = make_int4((int)c[0], (int)c[1], (int)c[2], 0);
        frame_s[s] = frame[j];
        jsel_s[s] = sel ? sel[j] : j;
        float r2 = 0.f;  // |k|^2 in the order of the tile spreader
        for (int a = 0; a < 3; ++a) r2 += kc[j * 3 + a] * kc[j * 3 + a];
        r2_s[s] = r2;
        npairs[s] = cnt;
    }
}

__global__ void bp3_emit_pairs_kernel(const int4* __restrict__ cell, const int* __restrict__ ppos,
                                      int64_t M, StreamGeom3 g, int* __restrict__ keys,
                                      int* __restrict__ vals) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < M;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int4 c = cell[s];
        int tx[SUPPORT_MAX], ty[SUPPORT_MAX], tz[SUPPORT_MAX];
        const int nx = support_blocks4(c.x, g.n[0], tx);
        const int ny = support_blocks4(c.y, g.n[1], ty);
        const int nz = support_blocks4(c.z, g.n[2], tz);
        int q = ppos[s];
        for (int cz = 0; cz < nz; ++cz)
            for (int b = 0; b < ny; ++b)
                for (int a = 0; a < nx; ++a) {
                    keys[q] = (tz[cz] * g.nb[1] + ty[b]) * g.nb[0] + tx[a];
                    vals[q] = (int)s;
                    ++q;
                }
    }
}

__global__ void bp3_block_count_kernel(const int* __restrict__ keys, int64_t n,
                                       int* __restrict__ counts) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&counts[keys[i]], 1);
}

// Per pair q (point s in block b): the kernel weights of s over the block's 4
// columns (pw[q][0..3]), rows (pw[q][4..7]) and planes (pw[q][8..11]), zero off
// the support and off the grid.
__global__ void bp3_pair_weights_kernel(const int* __restrict__ block_of,
                                        const float4* __restrict__ frac,
                                        const int4* __restrict__ cell, const int* __restrict__ pairs,
                                        int64_t npairs, StreamGeom3 g, float* __restrict__ pw) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < npairs;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int b = block_of[q];
        const int s = pairs[q];
        const int4 c = cell[s];
        const float4 f = frac[s];
        const int bx = b % g.nb[0], r = b / g.nb[0];
        const int by = r % g.nb[1], bz = r / g.nb[1];
        const int t0[3] = {bx * BP3_TB, by * BP3_TB, bz * BP3_TB};
        const int cc[3] = {c.x, c.y, c.z};
        const float ff[3] = {f.x, f.y, f.z};
        for (int a = 0; a < 3; ++a) {
            const int n = (int)g.n[a];
            for (int l = 0; l < BP3_TB; ++l) {
                float w = 0.f;
                const int x = t0[a] + l;
                if (x < n) {
                    int o = x - (cc[a] - 3);  // in (-n, n + 3): one wrap each way
                    o += o < 0 ? n : 0;
                    o -= o >= n ? n : 0;
                    if (o < 8) w = es_phi_f((float)(o - 3) - ff[a]);
                }
                pw[q * 12 + a * BP3_TB + l] = w;
            }
        }
    }
}

static StreamGeom3 geom3(const SpreadPlan& plan) {
    StreamGeom3 g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = plan.n[a];
        g.nb[a] = (int)ceil_div(plan.n[a], BP3_TB);
    }
    return g;
}

bool spread_bp3_supported(const msr_traj* traj) { return traj->M <= ((int64_t)INT32_MAX) / 64; }

void spread_streams3(const msr_traj* traj, SpreadPlan& plan, hipStream_t st) {
    if (plan.streams_ready) return;
    MSR_REQUIRE(plan.d == 3, MSR_ERR_ERROR, "3D output-block streams");
    const int64_t M = traj->M;
    // at most 4 blocks per dimension, 64 pairs per point: the int pair offsets
    // hold any total below 2^31
    MSR_REQUIRE(spread_bp3_supported(traj), MSR_ERR_UNSUPPORTED,
                "too many samples for the 3D backprojection streams");
    const StreamGeom3 g = geom3(plan);
    const int64_t nblocks = (int64_t)g.nb[0] * g.nb[1] * g.nb[2];
    MSR_REQUIRE(nblocks < ((int64_t)1 << 30), MSR_ERR_UNSUPPORTED, "grid too large");
    const int64_t Mp = std::max<int64_t>(M, 1);
    DevBuf frac;
    frac.alloc(sizeof(float4) * Mp);
    plan.cell.alloc(sizeof(int4) * Mp);
    plan.frame_s.alloc(sizeof(int) * Mp);
    plan.jsel_s.alloc(sizeof(int64_t) * Mp);
    plan.r2_s.alloc(sizeof(float) * Mp);
    DevBuf npp, ppos, keys, keys_sorted, vals, temp, counts;
    npp.alloc(sizeof(int) * (Mp + 1));
    ppos.alloc(sizeof(int) * (Mp + 1));
    const int grid = (int)std::min<int64_t>((M + 255) / 256, 8192);
    if (M > 0) {
        hipLaunchKernelGGL(bp3_points_kernel, dim3(grid), dim3(256), 0, st, traj->kc.as<float>(),
                           traj->frame.as<int>(), traj->sel.p ? traj->sel.as<int64_t>() : nullptr,
                           plan.perm.as<int>(), M, g, frac.as<float4>(), plan.cell.as<int4>(),
                           plan.frame_s.as<int>(), plan.jsel_s.as<int64_t>(),
                           plan.r2_s.as<float>(), npp.as<int>());
        MSR_CHECK_LAUNCH();
    }
    // pair offsets (exclusive scan over points, 64-bit total checked below)
    size_t tb = 0;
    MSR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, npp.as<int>(), ppos.as<int>(), (int)M + 1, st));
    temp.alloc(std::max<size_t>(tb, 16));
    fill_floats(npp.as<int>() + M, 1, 0.f, st);  // the scan's extra element: ppos[M] = total
    MSR_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb, npp.as<int>(), ppos.as<int>(), (int)M + 1, st));
    int total = 0;
    MSR_HIP(hipMemcpyAsync(&total, ppos.as<int>() + M, sizeof(int), hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    plan.npairs = total;
    const int64_t P = std::max<int64_t>(total, 1);
    keys.alloc(sizeof(int) * P);
    keys_sorted.alloc(sizeof(int) * P);
    vals.alloc(sizeof(int) * P);
    // zero pair records past the last pair: the gathers' prefetch reads up to
    // BP_PAD_PAIRS of them
    plan.pairs.alloc(sizeof(int) * (P + BP_PAD_PAIRS));
    fill_floats(plan.pairs.as<int>() + total, P - total + BP_PAD_PAIRS, 0.f, st);
    if (M > 0) {
        hipLaunchKernelGGL(bp3_emit_pairs_kernel, dim3(grid), dim3(256), 0, st, plan.cell.as<int4>(),
                           ppos.as<int>(), M, g, keys.as<int>(), vals.as<int>());
        MSR_CHECK_LAUNCH();
    }
    // stable sort by block: within a block the points stay in increasing order
    int bits = 1;
    while ((int64_t(1) << bits) < nblocks) ++bits;
    if (total > 0) {
        size_t sb = 0;
        MSR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sb, keys.as<int>(), keys_sorted.as<int>(),
                                                   vals.as<int>(), plan.pairs.as<int>(), total, 0,
                                                   bits, st));
        temp.alloc(std::max<size_t>(sb, 16));
        MSR_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, sb, keys.as<int>(), keys_sorted.as<int>(),
                                                   vals.as<int>(), plan.pairs.as<int>(), total, 0,
                                                   bits, st));
    }
    vals.release();
    keys.release();
    plan.pw.alloc(sizeof(float) * 12 * (P + BP_PAD_PAIRS));
    fill_floats(plan.pw.as<float>() + 12 * (int64_t)total, 12 * (P - total + BP_PAD_PAIRS), 0.f, st);
    if (total > 0) {
        const int g2 = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(bp3_pair_weights_kernel, dim3(g2), dim3(256), 0, st,
                           keys_sorted.as<int>(), frac.as<float4>(), plan.cell.as<int4>(),
                           plan.pairs.as<int>(), (int64_t)total, g, plan.pw.as<float>());
        MSR_CHECK_LAUNCH();
    }
    // per-block pair counts -> work items (host)
    counts.alloc(sizeof(int) * nblocks);
    fill_floats(counts.p, nblocks, 0.f, st);
    if (total > 0) {
        const int g2 = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(bp3_block_count_kernel, dim3(g2), dim3(256), 0, st,
                           keys_sorted.as<int>(), (int64_t)total, counts.as<int>());
        MSR_CHECK_LAUNCH();
    }
    std::vector<int> hc(nblocks);
    MSR_HIP(hipMemcpyAsync(hc.data(), counts.p, sizeof(int) * nblocks, hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    std::vector<int4> items;
    std::vector<int2> split;
    std::vector<int> split_n;
    int q = 0, slot = 0;
    for (int64_t t = 0; t < nblocks; ++t) {
        const int cnt = hc[t];
        if (cnt > ITEM_PAIRS) {
            split.push_back(make_int2((int)t, (int)items.size()));
            split_n.push_back((cnt + ITEM_PAIRS - 1) / ITEM_PAIRS);
            for (int q0 = q; q0 < q + cnt; q0 += ITEM_PAIRS)
                items.push_back(make_int4((int)t, q0, std::min(q + cnt, q0 + ITEM_PAIRS), slot++));
        } else if (cnt > 0) {
            items.push_back(make_int4((int)t, q, q + cnt, -1));
        }
        q += cnt;
    }
    plan.nitems = (int64_t)items.size();
    plan.nsplit = (int64_t)split.size();
    plan.items.alloc(sizeof(int4) * std::max<size_t>(items.size(), 1));
    plan.split_tiles.alloc(sizeof(int2) * std::max<size_t>(split.size(), 1));
    plan.split_nitems.alloc(sizeof(int) * std::max<size_t>(split.size(), 1));
    if (!items.empty())
        MSR_HIP(hipMemcpyAsync(plan.items.p, items.data(), sizeof(int4) * items.size(),
                               hipMemcpyHostToDevice, st));
    if (!split.empty()) {
        MSR_HIP(hipMemcpyAsync(plan.split_tiles.p, split.data(), sizeof(int2) * split.size(),
                               hipMemcpyHostToDevice, st));
        MSR_HIP(hipMemcpyAsync(plan.split_nitems.p, split_n.data(), sizeof(int) * split_n.size(),
                               hipMemcpyHostToDevice, st));
    }
    plan.nslots = slot;
    std::vector<uint64_t> key(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        const int b = items[i].x;
        const int c[3] = {b % g.nb[0], (b / g.nb[0]) % g.nb[1], b / (g.nb[0] * g.nb[1])};
        key[i] = bp_morton_key(c, 3);
    }
    const std::vector<int> order = bp_dispatch_order(items, key, plan.order_len);
    plan.order.alloc(sizeof(int) * order.size());
    MSR_HIP(hipMemcpyAsync(plan.order.p, order.data(), sizeof(int) * order.size(),
                           hipMemcpyHostToDevice, st));
    MSR_HIP(hipStreamSynchronize(st));  // host staging and temporaries go out of scope
    plan.streams_ready = true;
}

struct Bp3GatherArgs {
    const int4* items;
    const int* order;     // dispatch table (bp_dispatch_order)
    int64_t order_len;
    int ngp;              // channel groups in this launch
    const int* pairs;
    const float* pw;
    const float2* vals;   // (ngp, M, CG)
    int64_t M;
    int nch, g0;
    StreamGeom3 g;
    float2* grid;         // (nch, n2, n1, n0)
    float2* slab;         // (slots, nch, 64)
};

// One wavefront per (work item, channel group), lane (lx, ly, lz) on the cell
// of the 4 x 4 x 4 block; loads issued BP_DEPTH points ahead of their use, the
// pair word BP_DEPTH points before that (vector loads complete in order).
template <int CG>
__global__ void __launch_bounds__(64) bp3_gather_kernel(Bp3GatherArgs a) {
    constexpr int NV = CG == 1 ? 1 : 2 * CG / 16;  // value VGPRs per point
    constexpr int LV = CG == 1 ? 2 : 16;          // value floats per VGPR (per 16-lane row)
    constexpr int D = BP_DEPTH;
    int gl;
    const int ii = bp_item_of(a.order, a.order_len, a.ngp, gl);
    if (ii < 0) return;
    const int4 it = a.items[ii];
    const int blk = it.x, q0 = it.y, q1 = it.z, slot = it.w;
    const int ch0 = (a.g0 + gl) * CG;
    const int n0 = (int)a.g.n[0], n1 = (int)a.g.n[1], n2 = (int)a.g.n[2];
    const int bx = blk % a.g.nb[0], br = blk / a.g.nb[0];
    const int t0x = bx * BP3_TB, t0y = (br % a.g.nb[1]) * BP3_TB, t0z = (br / a.g.nb[1]) * BP3_TB;
    const int lane = threadIdx.x;
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    const float* __restrict__ vf =
        reinterpret_cast<const float*>(a.vals + (size_t)gl * a.M * CG) + (lane & 15) % LV;
    const float* __restrict__ pwx = a.pw + lx;
    const float* __restrict__ pwy = a.pw + BP3_TB + ly;
    const float* __restrict__ pwz = a.pw + 2 * BP3_TB + lz;
    const int* __restrict__ pq = a.pairs;
    float acc[2 * CG];
#pragma unroll
    for (int j = 0; j < 2 * CG; ++j) acc[j] = 0.f;
    const int n = q1 - q0;
    const int last = q1 - 1;
    __builtin_assume(n >= 1);
    int wd[D];               // pair words of points p + D .. p + 2D - 1
    float vv[D][NV];         // values of points p .. p + D - 1
    float wx[D], wy[D], wz[D];
    auto issue = [&](int i, int pt, int word) {  // loads of pair q0 + pt into stage i
        const int q = min(q0 + pt, last);
#pragma unroll
        for (int v = 0; v < NV; ++v) vv[i][v] = vf[(int64_t)word * (2 * CG) + 16 * v];
        wx[i] = pwx[(int64_t)q * 12];
        wy[i] = pwy[(int64_t)q * 12];
        wz[i] = pwz[(int64_t)q * 12];
    };
#pragma unroll
    for (int i = 0; i < D; ++i) issue(i, i, pq[min(q0 + i, last)]);
#pragma unroll
    for (int i = 0; i < D; ++i) wd[i] = pq[min(q0 + D + i, last)];
    asm volatile("" ::: "memory");  // prologue loads complete before the loop
    __builtin_amdgcn_s_waitcnt(BP_VMCNT0);
    for (int p = 0; p < n; p += D) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (p + i < n) {
                const float w = wx[i] * wy[i] * wz[i];
                if constexpr (CG == 1) {
                    fma_bcast2(acc, vv[i][0], w);
                } else {
                    fma_bcast16<0>(acc, vv[i][0], w);
                    if constexpr (NV > 1) fma_bcast16<16>(acc, vv[i][NV - 1], w);
                }
            }
            issue(i, p + i + D, wd[i]);
            __builtin_amdgcn_sched_barrier(0);  // next word into the same register
            wd[i] = pq[min(q0 + p + i + 2 * D, last)];
        }
    }
    const int x = t0x + lx, y = t0y + ly, z = t0z + lz;
    if (x >= n0 || y >= n1 || z >= n2) return;
    const int64_t G = a.g.n[0] * a.g.n[1] * a.g.n[2];
#pragma unroll
    for (int j = 0; j < CG; ++j) {
        if (ch0 + j >= a.nch) break;
        const float2 r = make_float2(acc[2 * j], acc[2 * j + 1]);
        if (slot < 0)
            a.grid[(int64_t)(ch0 + j) * G + ((int64_t)z * n1 + y) * n0 + x] = r;
        else
            a.slab[((int64_t)slot * a.nch + ch0 + j) * 64 + lane] = r;
    }
}

// kernel weight of cell (cx, cy, cz) (acc0) and (cx, cy, cz + 2) (acc1) of the
// 4 x 4 x 4 block, from the pair's weights over the block's x, y and z lines
struct Bp3Cells {
    static constexpr uint32_t PW_BYTES = 12 * sizeof(float);
    static constexpr int NW = 4;
    uint32_t ox, oy, oz;  // lane's byte offsets in a pair's weights: cx, cy, cz
    // weights of the pair whose record starts at p + rec
    __device__ __forceinline__ void load(const char* __restrict__ p, uint32_t rec,
                                         float (&w)[NW]) const {
        w[0] = bp_ldf(p, rec + ox);
        w[1] = bp_ldf(p, rec + oy);
        w[2] = bp_ldf(p, rec + oz);
        w[3] = bp_ldf(p + 2 * sizeof(float), rec + oz);
    }
    __device__ __forceinline__ void product(const float (&w)[NW], float& b0, float& b1) const {
        const float xy = w[0] * w[1];
        b0 = xy * w[2];
        b1 = xy * w[3];
    }
};

// The 16-channel group on the matrix cores (mfma_gather16): lane l ends with
// cell (l & 31) of z planes 0..1 in acc0 and cell (l & 31) + 32 in acc1, eight
// of the sixteen channels each (mfma_channel).
template <bool WIDE>
__global__ void __launch_bounds__(64) bp3_gather_mfma_kernel(Bp3GatherArgs a) {
    int gl;
    const int ii = bp_item_of(a.order, a.order_len, a.ngp, gl);
    if (ii < 0) return;
    const int4 it = a.items[ii];
    const int blk = it.x, q0 = it.y, q1 = it.z, slot = it.w;
    const int ch0 = (a.g0 + gl) * BP_CGMAX;
    const int n0 = (int)a.g.n[0], n1 = (int)a.g.n[1], n2 = (int)a.g.n[2];
    const int bx = blk % a.g.nb[0], br = blk / a.g.nb[0];
    const int t0x = bx * BP3_TB, t0y = (br % a.g.nb[1]) * BP3_TB, t0z = (br / a.g.nb[1]) * BP3_TB;
    const int lane = threadIdx.x;
    const int j = lane & 31, h = lane >> 5;
    const int cx = j & 3, cy = (j >> 2) & 3, cz = j >> 4;
    const float* __restrict__ vg = reinterpret_cast<const float*>(a.vals + (size_t)gl * a.M * BP_CGMAX);
    const Bp3Cells cells{(uint32_t)(cx * sizeof(float)), (uint32_t)((BP3_TB + cy) * sizeof(float)),
                         (uint32_t)((2 * BP3_TB + cz) * sizeof(float))};
    f32x16 acc0, acc1;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        acc0[r] = 0.f;
        acc1[r] = 0.f;
    }
    mfma_gather16<WIDE>(vg, a.pairs, a.pw, q0, q1, cells, acc0, acc1);
    const int64_t G = a.g.n[0] * a.g.n[1] * a.g.n[2];
    const int x = t0x + cx, y = t0y + cy;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const f32x16& acc = half ? acc1 : acc0;
        const int z = t0z + cz + 2 * half;
        if (x >= n0 || y >= n1 || z >= n2) continue;
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int c = ch0 + mfma_channel(r, h);
            if (c >= a.nch) continue;
            const float2 v = make_float2(acc[r], acc[r + 1]);
            if (slot < 0)
                a.grid[(int64_t)c * G + ((int64_t)z * n1 + y) * n0 + x] = v;
            else
                a.slab[((int64_t)slot * a.nch + c) * 64 + j + 32 * half] = v;
        }
    }
}

// split blocks: grid cells = sum of the block's partial blocks in item order
__global__ void bp3_split_reduce_kernel(const int2* __restrict__ split, const int* __restrict__ nit,
                                        const int4* __restrict__ items, const float2* __restrict__ slab,
                                        int nch, StreamGeom3 g, float2* __restrict__ grid) {
    const int si = blockIdx.x, ch = blockIdx.y;
    const int blk = split[si].x, first = split[si].y, cnt = nit[si];
    const int bx = blk % g.nb[0], br = blk / g.nb[0];
    const int t0x = bx * BP3_TB, t0y = (br % g.nb[1]) * BP3_TB, t0z = (br / g.nb[1]) * BP3_TB;
    const int64_t G = g.n[0] * g.n[1] * g.n[2];
    for (int r = threadIdx.x; r < 64; r += blockDim.x) {
        const int x = t0x + (r & 3), y = t0y + ((r >> 2) & 3), z = t0z + (r >> 4);
        if (x >= g.n[0] || y >= g.n[1] || z >= g.n[2]) continue;
        float2 acc = make_float2(0.f, 0.f);
        for (int i = 0; i < cnt; ++i) {
            const int s = items[first + i].w;
            const float2 v = slab[((int64_t)s * nch + ch) * 64 + r];
            acc.x += v.x;
            acc.y += v.y;
        }
        grid[(int64_t)ch * G + ((int64_t)z * g.n[1] + y) * g.n[0] + x] = acc;
    }
}

size_t spread_bp3_workspace_bytes(const msr_traj* traj, const SpreadPlan& plan, int nch,
                                  hipStream_t st) {
    spread_streams3(traj, const_cast<SpreadPlan&>(plan), st);
    const int CG = bp_group_size(nch);
    const int ng = (int)ceil_div(nch, CG);
    const int gp = bp_groups_per_pass(traj->M, CG, ng);
    return sizeof(float2) * ((size_t)plan.nslots * nch * 64 +
                             (size_t)gp * std::max<int64_t>(traj->M, 1) * CG);
}

void spread_bp3(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
                float2* grid, float2* work, hipStream_t st) {
    MSR_REQUIRE(plan.streams_ready && plan.d == 3, MSR_ERR_ERROR, "3D point streams not built");
    MSR_REQUIRE(src.kind == 0 && nch % src.K == 0, MSR_ERR_ERROR,
                "stream spreading takes backprojection values");
    const int64_t M = traj->M;
    const StreamGeom3 g = geom3(plan);
    const int64_t G = plan.n[0] * plan.n[1] * plan.n[2];
    fill_floats(grid, 2 * G * nch, 0.f, st);  // blocks no point reaches stay zero
    if (plan.nitems == 0) return;
    MSR_REQUIRE(plan.nitems < ((int64_t)1 << 31), MSR_ERR_UNSUPPORTED, "too many work items");
    const int CG = bp_group_size(nch);
    const int ng = (int)ceil_div(nch, CG);
    const int gp = bp_groups_per_pass(M, CG, ng);
    float2* slab = work;
    float2* vals = work + (size_t)plan.nslots * nch * 64;
    for (int g0 = 0; g0 < ng; g0 += gp) {
        const int ngp = std::min(gp, ng - g0);
        const int gr = bp_values_grid(M, CG);
#define MSR_BP3_VALUES(CGV)                                                                        \
    hipLaunchKernelGGL(bp_values_kernel<CGV>, dim3(gr, ngp), dim3(256), 0, st, src,                \
                       plan.jsel_s.as<int64_t>(), nullptr, plan.r2_s.as<float>(),                  \
                       plan.perm.as<int>(), plan.frame_s.as<int>(), M, nch, g0, vals)
        if (CG == 1) MSR_BP3_VALUES(1);
        else if (CG == 8) MSR_BP3_VALUES(8);
        else MSR_BP3_VALUES(BP_CGMAX);
#undef MSR_BP3_VALUES
        MSR_CHECK_LAUNCH();
        Bp3GatherArgs a{plan.items.as<int4>(), plan.order.as<int>(), plan.order_len, ngp,
                        plan.pairs.as<int>(), plan.pw.as<float>(), vals, M, nch, g0, g, grid, slab};
        const int64_t nwg = (int64_t)BP_XCDS * plan.order_len * ngp;
        MSR_REQUIRE(nwg < ((int64_t)1 << 31), MSR_ERR_UNSUPPORTED, "too many work items");
        dim3 grd((unsigned)nwg);
        if (CG == 1) hipLaunchKernelGGL(bp3_gather_kernel<1>, grd, dim3(64), 0, st, a);
        else if (CG == 8) hipLaunchKernelGGL(bp3_gather_kernel<8>, grd, dim3(64), 0, st, a);
        else if (!bp_wide_offsets(M))
            hipLaunchKernelGGL(bp3_gather_mfma_kernel<false>, grd, dim3(64), 0, st, a);
        else
            hipLaunchKernelGGL(bp3_gather_mfma_kernel<true>, grd, dim3(64), 0, st, a);
        MSR_CHECK_LAUNCH();
    }
    if (plan.nsplit > 0) {
        hipLaunchKernelGGL(bp3_split_reduce_kernel, dim3((unsigned)plan.nsplit, (unsigned)nch),
                           dim3(64), 0, st, plan.split_tiles.as<int2>(),
                           plan.split_nitems.as<int>(), plan.items.as<int4>(), slab, nch, g, grid);
        MSR_CHECK_LAUNCH();
    }
}

}  // namespace msr
