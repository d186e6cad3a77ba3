// 2D ComplexF32 spreading of the backprojection (C*K value channels; the
// reference runs one type-1 NUFFT per coil and coefficient).
//
// Output-owned blocks gathered in registers.  The grid is cut into 8 x 8
// blocks; every block has the list of the (tile-sorted) points whose W = 8
// support reaches it, in increasing point order, cut into work items of at
// most ITEM_PAIRS points.  One wavefront per (work item, group of CG
// channels) walks the item's points and sums w_x w_y v into the block's
// 64 cells.  Groups of 16 channels run on the matrix cores (mfma_gather16:
// the block's sums are the product of the points' values and their weights
// at the cells, two points per v_mfma_f32_32x32x2f32 pair).  Groups of 1 or 8
// channels keep one lane per cell: a point's 2 CG value floats are loaded 16
// per VGPR, the same 16 in each of the four 16-lane rows, and each FMA takes
// value j from lane j of its row through DPP row_newbcast, so a point costs
// 2 CG FMAs per cell and no other arithmetic.  A point's kernel weights over
// the block's columns and rows were laid out per (point, block) pair when the
// streams were built.  All loads are vector loads issued a few points ahead
// of their use (they complete in order).  Each cell's sum runs over the
// points in a fixed order (bit-reproducible; both forms give the same bits).  A block with one work item
// writes its cells straight into the grid; the items of a split block (the
// dense centre of a radial trajectory) write partial blocks that a second
// kernel sums in item order.
//
// Built once per trajectory and grid: each sorted point's cell, frame and
// data index, the pair lists (stable radix sort of (block, point) pairs by
// block) and the pair weights.  Per call: the channel values of every sorted
// point, conj(U[t, k]) y_c dcf, CG channels per point and group.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "spread_bp_common.hpp"

namespace msr {

constexpr int BP_TB = 8;           // block (cells per dimension): one lane per cell

void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

struct StreamGeom {
    int64_t n[2];
    int nb[2];
};

// distinct tiles (<= 3) holding the cells c-3 .. c+4 (periodic) of one dimension
__device__ __forceinline__ int support_tiles(int64_t c, int64_t n, int* t) {
    int cnt = 0;
    for (int o = -3; o <= 4; ++o) {
        int64_t a = c + o;
        if (a < 0) a += n;
        if (a >= n) a -= n;
        const int tt = (int)(a / BP_TB);
        bool seen = false;
        for (int i = 0; i < cnt; ++i) seen |= (t[i] == tt);
        if (!seen) t[cnt++] = tt;
    }
    return cnt;
}

__global__ void bp_points_kernel(const float* __restrict__ kc, const int* __restrict__ frame,
                                 const int64_t* __restrict__ sel, const int* __restrict__ perm,
                                 int64_t M, StreamGeom g, float2* __restrict__ frac,
                                 int2* __restrict__ cell, int* __restrict__ frame_s,
                                 int64_t* __restrict__ jsel_s, float2* __restrict__ kc_s,
                                 int* __restrict__ npairs) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < M;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[s];
        int64_t c[2];
        float f[2];
        for (int a = 0; a < 2; ++a) {
            const double x = (double)kc[j * 2 + a] * (double)g.n[a];
            const double fl = floor(x);
            f[a] = (float)(x - fl);
            int64_t cc = (int64_t)fl % g.n[a];
            if (cc < 0) cc += g.n[a];
            c[a] = cc;
        }
        frac[s] = make_float2(f[0], f[1]);
        cell[s] = make_int2((int)c[0], (int)c[1]);
        frame_s[s] = frame[j];
        jsel_s[s] = sel ? sel[j] : j;
        kc_s[s] = make_float2(kc[j * 2], kc[j * 2 + 1]);
        int tx[3], ty[3];
        npairs[s] = support_tiles(c[0], g.n[0], tx) * support_tiles(c[1], g.n[1], ty);
    }
}

__global__ void bp_emit_pairs_kernel(const int2* __restrict__ cell, const int* __restrict__ ppos,
                                     int64_t M, StreamGeom g, int* __restrict__ keys,
                                     int* __restrict__ vals) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < M;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int2 c = cell[s];
        int tx[3], ty[3];
        const int nx = support_tiles(c.x, g.n[0], tx);
        const int ny = support_tiles(c.y, g.n[1], ty);
        int q = ppos[s];
        for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
                keys[q] = ty[b] * g.nb[0] + tx[a];
                vals[q] = (int)s;
                ++q;
            }
    }
}

__global__ void bp_tile_count_kernel(const int* __restrict__ keys, int64_t n, int* __restrict__ counts) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&counts[keys[i]], 1);
}

// Per pair q (point s in block t): the kernel weights of s over the block's 8
// columns (pw[q][0..7]) and 8 rows (pw[q][8..15]), zero off the support and
// off the grid.
__global__ void bp_pair_weights_kernel(const int* __restrict__ tile_of, const float2* __restrict__ frac,
                                       const int2* __restrict__ cell, const int* __restrict__ pairs,
                                       int64_t npairs, StreamGeom g, float* __restrict__ pw) {
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < npairs;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int tile = tile_of[q];
        const int s = pairs[q];
        const int2 c = cell[s];
        const float2 f = frac[s];
        const int t0[2] = {(tile % g.nb[0]) * BP_TB, (tile / g.nb[0]) * BP_TB};
        const int cc[2] = {c.x, c.y};
        const float ff[2] = {f.x, f.y};
        for (int a = 0; a < 2; ++a) {
            const int n = (int)g.n[a];
            for (int l = 0; l < BP_TB; ++l) {
                float w = 0.f;
                const int x = t0[a] + l;
                if (x < n) {
                    int o = x - (cc[a] - 3);  // in (-n, n + 3): one wrap each way
                    o += o < 0 ? n : 0;
                    o -= o >= n ? n : 0;
                    if (o < 8) w = es_phi_f((float)(o - 3) - ff[a]);
                }
                pw[q * 16 + a * BP_TB + l] = w;
            }
        }
    }
}

void spread_streams(const msr_traj* traj, SpreadPlan& plan, hipStream_t st) {
    if (plan.streams_ready) return;
    MSR_REQUIRE(plan.d == 2, MSR_ERR_ERROR, "output-tile streams are 2D");
    const int64_t M = traj->M;
    MSR_REQUIRE(M < ((int64_t)1 << 30), MSR_ERR_UNSUPPORTED, "more than 2^30 samples");
    StreamGeom g;
    for (int a = 0; a < 2; ++a) {
        g.n[a] = plan.n[a];
        g.nb[a] = (int)ceil_div(plan.n[a], BP_TB);
    }
    const int64_t ntiles = (int64_t)g.nb[0] * g.nb[1];
    const int64_t Mp = std::max<int64_t>(M, 1);
    DevBuf frac;
    frac.alloc(sizeof(float2) * Mp);
    plan.cell.alloc(sizeof(int2) * Mp);
    plan.frame_s.alloc(sizeof(int) * Mp);
    plan.jsel_s.alloc(sizeof(int64_t) * Mp);
    plan.kc_s.alloc(sizeof(float2) * Mp);
    DevBuf npp, ppos, keys, keys_sorted, vals, temp, counts;
    npp.alloc(sizeof(int) * (Mp + 1));
    ppos.alloc(sizeof(int) * (Mp + 1));
    const int grid = (int)std::min<int64_t>((M + 255) / 256, 8192);
    if (M > 0) {
        hipLaunchKernelGGL(bp_points_kernel, dim3(grid), dim3(256), 0, st, traj->kc.as<float>(),
                           traj->frame.as<int>(), traj->sel.p ? traj->sel.as<int64_t>() : nullptr,
                           plan.perm.as<int>(), M, g, frac.as<float2>(), plan.cell.as<int2>(),
                           plan.frame_s.as<int>(), plan.jsel_s.as<int64_t>(),
                           plan.kc_s.as<float2>(), npp.as<int>());
        MSR_CHECK_LAUNCH();
    }
    // pair offsets (exclusive scan over points) and the total
    size_t tb = 0;
    MSR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, npp.as<int>(), ppos.as<int>(), (int)M + 1, st));
    temp.alloc(std::max<size_t>(tb, 16));
    // npp[M] is the scan's extra element: zero it so ppos[M] is the total
    fill_floats(npp.as<int>() + M, 1, 0.f, st);
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
        hipLaunchKernelGGL(bp_emit_pairs_kernel, dim3(grid), dim3(256), 0, st, plan.cell.as<int2>(),
                           ppos.as<int>(), M, g, keys.as<int>(), vals.as<int>());
        MSR_CHECK_LAUNCH();
    }
    // stable sort by tile: within a tile the points stay in increasing order
    int bits = 1;
    while ((int64_t(1) << bits) < ntiles) ++bits;
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
    // pair weights (keys_sorted: the block of each pair)
    plan.pw.alloc(sizeof(float) * 16 * (P + BP_PAD_PAIRS));
    fill_floats(plan.pw.as<float>() + 16 * (int64_t)total, 16 * (P - total + BP_PAD_PAIRS), 0.f, st);
    if (total > 0) {
        const int g2 = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(bp_pair_weights_kernel, dim3(g2), dim3(256), 0, st, keys_sorted.as<int>(),
                           frac.as<float2>(), plan.cell.as<int2>(), plan.pairs.as<int>(),
                           (int64_t)total, g, plan.pw.as<float>());
        MSR_CHECK_LAUNCH();
    }
    // per-tile pair counts -> work items (host)
    counts.alloc(sizeof(int) * ntiles);
    fill_floats(counts.p, ntiles, 0.f, st);
    if (total > 0) {
        const int g2 = (int)std::min<int64_t>((total + 255) / 256, 8192);
        hipLaunchKernelGGL(bp_tile_count_kernel, dim3(g2), dim3(256), 0, st, keys_sorted.as<int>(),
                           (int64_t)total, counts.as<int>());
        MSR_CHECK_LAUNCH();
    }
    std::vector<int> hc(ntiles);
    MSR_HIP(hipMemcpyAsync(hc.data(), counts.p, sizeof(int) * ntiles, hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    std::vector<int4> items;
    std::vector<int2> split;
    std::vector<int> split_n;
    int q = 0, slot = 0;
    for (int64_t t = 0; t < ntiles; ++t) {
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
        const int c[2] = {items[i].x % g.nb[0], items[i].x / g.nb[0]};
        key[i] = bp_morton_key(c, 2);
    }
    const std::vector<int> order = bp_dispatch_order(items, key, plan.order_len);
    plan.order.alloc(sizeof(int) * order.size());
    MSR_HIP(hipMemcpyAsync(plan.order.p, order.data(), sizeof(int) * order.size(),
                           hipMemcpyHostToDevice, st));
    MSR_HIP(hipStreamSynchronize(st));  // host staging and temporaries go out of scope
    plan.streams_ready = true;
}

struct BpGatherArgs {
    const int4* items;
    const int* order;     // dispatch table (bp_dispatch_order)
    int64_t order_len;
    int ngp;              // channel groups in this launch
    const int* pairs;
    const float* pw;
    const float2* vals;   // (ngp, M, CG)
    int64_t M;
    int nch, g0;
    StreamGeom g;
    float2* grid;         // (nch, n1, n0)
    float2* slab;         // (slots, nch, TB*TB)
};


// One wavefront per (work item, channel group), lane (lx, ly) on the cell
// (lx, ly) of the 8 x 8 block.  The loads of a point (its 2 CG value floats,
// 16 per VGPR replicated over the four rows, and the lane's two weights) are
// issued BP_DEPTH points ahead of its use, the pair word (point index) BP_DEPTH
// points before that; vector loads complete in order, so each use waits only
// for loads issued BP_DEPTH points earlier.
template <int CG>
__global__ void __launch_bounds__(64) bp_gather_kernel(BpGatherArgs a) {
    constexpr int NV = CG == 1 ? 1 : 2 * CG / 16;  // value VGPRs per point
    constexpr int LV = CG == 1 ? 2 : 16;          // value floats per VGPR (per 16-lane row)
    constexpr int D = BP_DEPTH;
    int gl;
    const int ii = bp_item_of(a.order, a.order_len, a.ngp, gl);
    if (ii < 0) return;
    const int4 it = a.items[ii];
    const int tile = it.x, q0 = it.y, q1 = it.z, slot = it.w;
    const int ch0 = (a.g0 + gl) * CG;
    const int n0 = (int)a.g.n[0], n1 = (int)a.g.n[1];
    const int t0x = (tile % a.g.nb[0]) * BP_TB, t0y = (tile / a.g.nb[0]) * BP_TB;
    const int lane = threadIdx.x;
    const int lx = lane & 7, ly = lane >> 3;
    const float* __restrict__ vf =
        reinterpret_cast<const float*>(a.vals + (size_t)gl * a.M * CG) + (lane & 15) % LV;
    const float* __restrict__ pwx = a.pw + lx;
    const float* __restrict__ pwy = a.pw + BP_TB + ly;
    const int* __restrict__ pq = a.pairs;
    float acc[2 * CG];
#pragma unroll
    for (int j = 0; j < 2 * CG; ++j) acc[j] = 0.f;
    const int n = q1 - q0;
    const int last = q1 - 1;
    __builtin_assume(n >= 1);
    int wd[D];            // pair words of points p + D .. p + 2D - 1
    float vv[D][NV];      // values of points p .. p + D - 1
    float wx[D], wy[D];   // weights of points p .. p + D - 1
    auto issue = [&](int i, int pt, int word) {  // loads of point q0 + pt into stage i
        const int q = min(q0 + pt, last);
#pragma unroll
        for (int v = 0; v < NV; ++v) vv[i][v] = vf[(int64_t)word * (2 * CG) + 16 * v];
        wx[i] = pwx[(int64_t)q * 16];
        wy[i] = pwy[(int64_t)q * 16];
    };
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const int w0 = pq[min(q0 + i, last)];
        issue(i, i, w0);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) wd[i] = pq[min(q0 + D + i, last)];
    asm volatile("" ::: "memory");  // prologue loads complete before the loop
    __builtin_amdgcn_s_waitcnt(BP_VMCNT0);
    for (int p = 0; p < n; p += D) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (p + i < n) {
                const float w = wx[i] * wy[i];
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
    const int ex = min(BP_TB, n0 - t0x), ey = min(BP_TB, n1 - t0y);
    if (lx >= ex || ly >= ey) return;
    const int64_t G = a.g.n[0] * a.g.n[1];
#pragma unroll
    for (int j = 0; j < CG; ++j) {
        if (ch0 + j >= a.nch) break;
        const float2 r = make_float2(acc[2 * j], acc[2 * j + 1]);
        if (slot < 0)
            a.grid[(int64_t)(ch0 + j) * G + (int64_t)(t0y + ly) * n0 + t0x + lx] = r;
        else
            a.slab[((int64_t)slot * a.nch + ch0 + j) * (BP_TB * BP_TB) + lane] = r;
    }
    (void)n1;
}

// kernel weight of cell (cx, cy) (acc0) and (cx, cy + 4) (acc1) of the 8 x 8
// block, from the pair's weights over the block's columns and rows
struct Bp2Cells {
    static constexpr uint32_t PW_BYTES = 16 * sizeof(float);
    static constexpr int NW = 3;
    uint32_t ox, oy;  // lane's byte offsets in a pair's weights: column cx, row cy
    // weights of the pair whose record starts at p + rec
    __device__ __forceinline__ void load(const char* __restrict__ p, uint32_t rec,
                                         float (&w)[NW]) const {
        w[0] = bp_ldf(p, rec + ox);
        w[1] = bp_ldf(p, rec + oy);
        w[2] = bp_ldf(p + 4 * sizeof(float), rec + oy);
    }
    __device__ __forceinline__ void product(const float (&w)[NW], float& b0, float& b1) const {
        b0 = w[0] * w[1];
        b1 = w[0] * w[2];
    }
};

// The 16-channel group on the matrix cores (mfma_gather16): one wavefront per
// (work item, group); lane l ends with cell (l & 31) of rows 0..3 in acc0 and
// cell (l & 31) + 32 in acc1, eight of the sixteen channels each
// (mfma_channel).
template <bool WIDE>
__global__ void __launch_bounds__(64) bp_gather_mfma_kernel(BpGatherArgs a) {
    int gl;
    const int ii = bp_item_of(a.order, a.order_len, a.ngp, gl);
    if (ii < 0) return;
    const int4 it = a.items[ii];
    const int tile = it.x, q0 = it.y, q1 = it.z, slot = it.w;
    const int ch0 = (a.g0 + gl) * BP_CGMAX;
    const int n0 = (int)a.g.n[0], n1 = (int)a.g.n[1];
    const int t0x = (tile % a.g.nb[0]) * BP_TB, t0y = (tile / a.g.nb[0]) * BP_TB;
    const int lane = threadIdx.x;
    const int j = lane & 31, h = lane >> 5;
    const int cx = j & 7, cy = j >> 3;
    const float* __restrict__ vg = reinterpret_cast<const float*>(a.vals + (size_t)gl * a.M * BP_CGMAX);
    const Bp2Cells cells{(uint32_t)(cx * sizeof(float)), (uint32_t)((BP_TB + cy) * sizeof(float))};
    f32x16 acc0, acc1;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        acc0[r] = 0.f;
        acc1[r] = 0.f;
    }
    mfma_gather16<WIDE>(vg, a.pairs, a.pw, q0, q1, cells, acc0, acc1);
    const int ex = min(BP_TB, n0 - t0x), ey = min(BP_TB, n1 - t0y);
    const int64_t G = a.g.n[0] * a.g.n[1];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        const f32x16& acc = half ? acc1 : acc0;
        const int ly = cy + 4 * half;
        if (cx >= ex || ly >= ey) continue;
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int c = ch0 + mfma_channel(r, h);
            if (c >= a.nch) continue;
            const float2 v = make_float2(acc[r], acc[r + 1]);
            if (slot < 0)
                a.grid[(int64_t)c * G + (int64_t)(t0y + ly) * n0 + t0x + cx] = v;
            else
                a.slab[((int64_t)slot * a.nch + c) * (BP_TB * BP_TB) + ly * BP_TB + cx] = v;
        }
    }
    (void)n1;
}

// split blocks: grid cells = sum of the block's partial blocks in item order
__global__ void bp_split_reduce_kernel(const int2* __restrict__ split, const int* __restrict__ nit,
                                       const int4* __restrict__ items, const float2* __restrict__ slab,
                                       int nch, StreamGeom g, float2* __restrict__ grid) {
    const int si = blockIdx.x, ch = blockIdx.y;
    const int tile = split[si].x, first = split[si].y, cnt = nit[si];
    const int t0x = (tile % g.nb[0]) * BP_TB, t0y = (tile / g.nb[0]) * BP_TB;
    const int ex = min(BP_TB, (int)g.n[0] - t0x);
    const int ey = min(BP_TB, (int)g.n[1] - t0y);
    const int64_t G = g.n[0] * g.n[1];
    for (int r = threadIdx.x; r < BP_TB * BP_TB; r += blockDim.x) {
        const int ry = r / BP_TB, rx = r - ry * BP_TB;
        if (rx >= ex || ry >= ey) continue;
        float2 acc = make_float2(0.f, 0.f);
        for (int i = 0; i < cnt; ++i) {
            const int s = items[first + i].w;
            const float2 v = slab[((int64_t)s * nch + ch) * (BP_TB * BP_TB) + r];
            acc.x += v.x;
            acc.y += v.y;
        }
        grid[(int64_t)ch * G + (int64_t)(t0y + ry) * g.n[0] + t0x + rx] = acc;
    }
}

size_t spread_bp_workspace_bytes(const msr_traj* traj, const SpreadPlan& plan, int nch, int ncoils,
                                 hipStream_t st) {
    (void)ncoils;
    spread_streams(traj, const_cast<SpreadPlan&>(plan), st);
    const int CG = bp_group_size(nch);
    const int ng = (int)ceil_div(nch, CG);
    const int gp = bp_groups_per_pass(traj->M, CG, ng);
    return sizeof(float2) * ((size_t)plan.nslots * nch * BP_TB * BP_TB +
                             (size_t)gp * std::max<int64_t>(traj->M, 1) * CG);
}

void spread_bp(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
               float2* grid, float2* work, hipStream_t st) {
    MSR_REQUIRE(plan.streams_ready, MSR_ERR_ERROR, "point streams not built");
    MSR_REQUIRE(src.kind == 0 && nch % src.K == 0, MSR_ERR_ERROR,
                "stream spreading takes backprojection values");
    const int64_t M = traj->M;
    const int64_t G = plan.n[0] * plan.n[1];
    StreamGeom g;
    for (int a = 0; a < 2; ++a) {
        g.n[a] = plan.n[a];
        g.nb[a] = (int)ceil_div(plan.n[a], BP_TB);
    }
    // tiles no point reaches stay zero
    fill_floats(grid, 2 * G * nch, 0.f, st);
    if (plan.nitems == 0) return;
    MSR_REQUIRE(plan.nitems < ((int64_t)1 << 31), MSR_ERR_UNSUPPORTED, "too many work items");
    const int CG = bp_group_size(nch);
    const int ng = (int)ceil_div(nch, CG);
    const int gp = bp_groups_per_pass(M, CG, ng);
    float2* slab = work;
    float2* vals = work + (size_t)plan.nslots * nch * BP_TB * BP_TB;
    for (int g0 = 0; g0 < ng; g0 += gp) {
        const int ngp = std::min(gp, ng - g0);
        const int gr = bp_values_grid(M, CG);
#define MSR_BP_VALUES(CGV)                                                                         \
    hipLaunchKernelGGL(bp_values_kernel<CGV>, dim3(gr, ngp), dim3(256), 0, st, src,                \
                       plan.jsel_s.as<int64_t>(), plan.kc_s.as<float2>(), nullptr,                 \
                       plan.perm.as<int>(),                                                        \
                       plan.frame_s.as<int>(), M, nch, g0, vals)
        if (CG == 1) MSR_BP_VALUES(1);
        else if (CG == 8) MSR_BP_VALUES(8);
        else MSR_BP_VALUES(BP_CGMAX);
#undef MSR_BP_VALUES
        MSR_CHECK_LAUNCH();
        BpGatherArgs a{plan.items.as<int4>(), plan.order.as<int>(), plan.order_len, ngp,
                       plan.pairs.as<int>(), plan.pw.as<float>(), vals, M, nch, g0, g, grid, slab};
        const int64_t nwg = (int64_t)BP_XCDS * plan.order_len * ngp;
        MSR_REQUIRE(nwg < ((int64_t)1 << 31), MSR_ERR_UNSUPPORTED, "too many work items");
        dim3 grd((unsigned)nwg);
        if (CG == 1) hipLaunchKernelGGL(bp_gather_kernel<1>, grd, dim3(64), 0, st, a);
        else if (CG == 8) hipLaunchKernelGGL(bp_gather_kernel<8>, grd, dim3(64), 0, st, a);
        else if (!bp_wide_offsets(M))
            hipLaunchKernelGGL(bp_gather_mfma_kernel<false>, grd, dim3(64), 0, st, a);
        else
            hipLaunchKernelGGL(bp_gather_mfma_kernel<true>, grd, dim3(64), 0, st, a);
        MSR_CHECK_LAUNCH();
    }
    if (plan.nsplit > 0) {
        hipLaunchKernelGGL(bp_split_reduce_kernel, dim3((unsigned)plan.nsplit, (unsigned)nch),
                           dim3(256), 0, st, plan.split_tiles.as<int2>(),
                           plan.split_nitems.as<int>(), plan.items.as<int4>(), slab, nch, g, grid);
        MSR_CHECK_LAUNCH();
    }
}

}  // namespace msr
