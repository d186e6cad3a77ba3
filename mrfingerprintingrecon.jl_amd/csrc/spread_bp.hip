// 2D ComplexF32 spreading of the backprojection (C*K value channels; the
// reference runs one type-1 NUFFT per coil and coefficient).
//
// Output-owned tiles.  The grid is cut into 16 x 16 tiles; every tile has the
// list of the (tile-sorted) points whose W = 8 support reaches it, in
// increasing point order, cut into work items of at most ITEM_PAIRS points.
// One wavefront per (work item, group of BP_CH channels): it walks the item's
// points and lane (ox, oy) adds w_x[ox] w_y[oy] v to its cell, when that cell
// lies in the tile, in each of the group's channel tiles (plain LDS
// read-modify-writes in program order: the accumulation order is fixed and the
// result bit-reproducible; the BP_CH updates per point go to different tiles
// and overlap in the LDS pipeline).  A tile with one work item writes its
// cells straight into the grid; the items of a split tile (the dense centre
// of a radial trajectory) write partial tiles that a second kernel sums in
// item order.  Points are staged 64 at a time in LDS (lane = point).
//
// Built once per trajectory and grid: each sorted point's kernel weights,
// cell, frame and data index, and the pair lists (stable radix sort of
// (tile, point) pairs by tile).  Per call: the coil data in sorted point order
// with the density compensation applied.
#include <algorithm>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "spread.hpp"

namespace msr {

constexpr int ITEM_PAIRS = 4096;
constexpr int BP_TB = 16;          // tile (cells per dimension)
constexpr int BP_PITCH = 24;       // LDS row pitch: the 8 x 8 lanes of an update hit 64 banks
constexpr int BP_CH = 4;           // channels per wavefront (one wavefront per workgroup)
constexpr int BP_CHUNK = 64;

void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

__device__ __forceinline__ float es_phi_f(float t) {
    const float z = t * (2.0f / SPREAD_W);
    const float a = 1.0f - z * z;
    return a > 0.0f ? expf(SPREAD_BETA * (sqrtf(a) - 1.0f)) : 0.0f;
}

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
                                 int64_t M, StreamGeom g, float* __restrict__ wts,
                                 int2* __restrict__ cell, int* __restrict__ frame_s,
                                 int64_t* __restrict__ jsel_s, float2* __restrict__ kc_s,
                                 int* __restrict__ npairs) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < M;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[s];
        int64_t c[2];
        for (int a = 0; a < 2; ++a) {
            const double x = (double)kc[j * 2 + a] * (double)g.n[a];
            const double fl = floor(x);
            const float f = (float)(x - fl);
            int64_t cc = (int64_t)fl % g.n[a];
            if (cc < 0) cc += g.n[a];
            c[a] = cc;
            for (int o = 0; o < 8; ++o) wts[s * 16 + a * 8 + o] = es_phi_f((float)(o - 3) - f);
        }
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

void spread_streams(const msr_traj* traj, SpreadPlan& plan, hipStream_t st) {
    if (plan.streams_ready) return;
    MSR_REQUIRE(plan.d == 2, MSR_ERR_ERROR, "output-tile streams are 2D");
    const int64_t M = traj->M;
    MSR_REQUIRE(M < ((int64_t)1 << 28), MSR_ERR_UNSUPPORTED, "more than 2^28 samples");
    StreamGeom g;
    for (int a = 0; a < 2; ++a) {
        g.n[a] = plan.n[a];
        g.nb[a] = (int)ceil_div(plan.n[a], BP_TB);
    }
    const int64_t ntiles = (int64_t)g.nb[0] * g.nb[1];
    const int64_t Mp = std::max<int64_t>(M, 1);
    plan.wts.alloc(sizeof(float) * 16 * Mp);
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
                           plan.perm.as<int>(), M, g, plan.wts.as<float>(), plan.cell.as<int2>(),
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
    plan.pairs.alloc(sizeof(int) * P);
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
    MSR_HIP(hipStreamSynchronize(st));  // host staging and temporaries go out of scope
    plan.streams_ready = true;
}

// ------------------------------------------------------------------ per call
// ys[c][s] = y[jsel(s), c] * dcf(s): coil data in sorted point order
__global__ void bp_sorted_data_kernel(ValueSrc src, const int64_t* __restrict__ jsel_s,
                                      const float2* __restrict__ kc_s, const int* __restrict__ perm,
                                      int64_t M, int C, float2* __restrict__ ys) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < M * C;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(e / M);
        const int64_t s = e - (int64_t)c * M;
        float2 v = src.data[jsel_s[s] + (int64_t)c * src.nsel];
        if (src.dcf_mode == 1 || src.dcf_mode == 2) {
            const float2 q = kc_s[s];
            const float r2 = q.x * q.x + q.y * q.y;
            v = cscale(v, src.dcf_mode == 1 ? sqrtf(r2) : r2);
        } else if (src.dcf_mode == 3) {
            v = cscale(v, src.dcf[perm[s]]);
        }
        ys[e] = v;
    }
}

__device__ __forceinline__ float2 basis_val(const ValueSrc& s, int t, int k) {
    if (!s.U) return make_float2(t == k ? 1.f : 0.f, 0.f);
    if (s.u_real) return make_float2(reinterpret_cast<const float*>(s.U)[t + (int64_t)k * s.nt], 0.f);
    return reinterpret_cast<const float2*>(s.U)[t + (int64_t)k * s.nt];
}

struct BpTileArgs {
    const int4* items;
    const int* pairs;
    const float* wts;
    const int2* cell;
    const int* frame_s;
    const float2* ys;     // (C, M)
    int64_t M;
    int nch;
    StreamGeom g;
    float2* grid;         // (nch, n1, n0)
    float2* slab;         // (slots, nch, TB*TB)
};

// One wavefront per workgroup, BP_CH channels per wavefront: for every point
// the BP_CH read-modify-writes go to different tiles, so they overlap in the
// LDS pipeline; the order per tile is the program order (deterministic).
__global__ void __launch_bounds__(64) bp_tile_kernel(BpTileArgs a, ValueSrc src) {
    __shared__ float tre[BP_CH][BP_TB * BP_PITCH];
    __shared__ float tim[BP_CH][BP_TB * BP_PITCH];
    __shared__ float2 sv[BP_CH][BP_CHUNK];
    __shared__ int2 scell[BP_CHUNK];
    __shared__ float sw[BP_CHUNK][16];
    const int4 it = a.items[blockIdx.x];
    const int tile = it.x, q0 = it.y, q1 = it.z, slot = it.w;
    const int ch0 = blockIdx.y * BP_CH;
    const int n0 = (int)a.g.n[0], n1 = (int)a.g.n[1];
    const int t0x = (tile % a.g.nb[0]) * BP_TB, t0y = (tile / a.g.nb[0]) * BP_TB;
    const int ex = min(BP_TB, n0 - t0x);
    const int ey = min(BP_TB, n1 - t0y);
    const int lane = threadIdx.x;
    for (int e = lane; e < BP_CH * BP_TB * BP_PITCH; e += 64) {
        (&tre[0][0])[e] = 0.f;
        (&tim[0][0])[e] = 0.f;
    }
    const int ox = lane & 7, oy = lane >> 3;
    int cc[BP_CH], kk[BP_CH];
#pragma unroll
    for (int h = 0; h < BP_CH; ++h) {
        const int ch = min(ch0 + h, a.nch - 1);
        cc[h] = ch / src.K;
        kk[h] = ch - cc[h] * src.K;
    }
    const int nchunk = (q1 - q0 + BP_CHUNK - 1) / BP_CHUNK;
    for (int n = 0; n < nchunk; ++n) {
        // stage the chunk: lane = point
        const int q = q0 + n * BP_CHUNK + lane;
        const int np = min(BP_CHUNK, q1 - q0 - n * BP_CHUNK);
        if (lane < np) {
            const int s = a.pairs[q];
            const int t = a.frame_s[s];
#pragma unroll
            for (int h = 0; h < BP_CH; ++h)
                sv[h][lane] = cmulc(basis_val(src, t, kk[h]), a.ys[(int64_t)cc[h] * a.M + s]);
            scell[lane] = a.cell[s];
            const float4* w4 = reinterpret_cast<const float4*>(a.wts + (int64_t)s * 16);
#pragma unroll
            for (int r = 0; r < 4; ++r) reinterpret_cast<float4*>(sw[lane])[r] = w4[r];
        }
        __syncthreads();
        // software pipeline: the staged data of point p + 1 is read while
        // point p's read-modify-writes are in flight
        int2 cn = scell[0];
        float wxn = sw[0][ox], wyn = sw[0][8 + oy];
        float2 vn[BP_CH];
#pragma unroll
        for (int h = 0; h < BP_CH; ++h) vn[h] = sv[h][0];
        for (int p = 0; p < np; ++p) {
            const int2 c = cn;
            const float w = wxn * wyn;
            float2 v[BP_CH];
#pragma unroll
            for (int h = 0; h < BP_CH; ++h) v[h] = vn[h];
            if (p + 1 < np) {
                cn = scell[p + 1];
                wxn = sw[p + 1][ox];
                wyn = sw[p + 1][8 + oy];
#pragma unroll
                for (int h = 0; h < BP_CH; ++h) vn[h] = sv[h][p + 1];
            }
            int ax = c.x - 3 + ox;
            ax += ax < 0 ? n0 : 0;
            ax -= ax >= n0 ? n0 : 0;
            int ay = c.y - 3 + oy;
            ay += ay < 0 ? n1 : 0;
            ay -= ay >= n1 ? n1 : 0;
            const int rx = ax - t0x, ry = ay - t0y;
            if (rx >= 0 && rx < ex && ry >= 0 && ry < ey) {
                const int o = ry * BP_PITCH + rx;
#pragma unroll
                for (int h = 0; h < BP_CH; ++h) {
                    tre[h][o] += w * v[h].x;
                    tim[h][o] += w * v[h].y;
                }
            }
        }
        __syncthreads();
    }
    // output: the tile's cells into the grid, or a partial tile into the slab
    const int chg = min(BP_CH, a.nch - ch0);
    const int64_t G = a.g.n[0] * a.g.n[1];
    for (int e = lane; e < chg * BP_TB * BP_TB; e += 64) {
        const int k = e / (BP_TB * BP_TB), r = e - k * (BP_TB * BP_TB);
        const int ry = r / BP_TB, rx = r - ry * BP_TB;
        const float2 v = make_float2(tre[k][ry * BP_PITCH + rx], tim[k][ry * BP_PITCH + rx]);
        if (slot < 0) {
            if (rx < ex && ry < ey)
                a.grid[(int64_t)(ch0 + k) * G + (int64_t)(t0y + ry) * a.g.n[0] + t0x + rx] = v;
        } else {
            a.slab[((int64_t)slot * a.nch + ch0 + k) * (BP_TB * BP_TB) + r] = v;
        }
    }
}

// split tiles: grid cells = sum of the tile's partial tiles in item order
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
    spread_streams(traj, const_cast<SpreadPlan&>(plan), st);
    return sizeof(float2) * ((size_t)plan.nslots * nch * BP_TB * BP_TB +
                             (size_t)ncoils * std::max<int64_t>(traj->M, 1));
}

void spread_bp(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
               float2* grid, float2* work, hipStream_t st) {
    MSR_REQUIRE(plan.streams_ready, MSR_ERR_ERROR, "point streams not built");
    MSR_REQUIRE(src.kind == 0 && nch % src.K == 0, MSR_ERR_ERROR,
                "stream spreading takes backprojection values");
    const int C = nch / src.K;
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
    float2* slab = work;
    float2* ys = work + (size_t)plan.nslots * nch * BP_TB * BP_TB;
    const int gr = (int)std::min<int64_t>((M * C + 255) / 256, 16384);
    hipLaunchKernelGGL(bp_sorted_data_kernel, dim3(gr), dim3(256), 0, st, src,
                       plan.jsel_s.as<int64_t>(), plan.kc_s.as<float2>(), plan.perm.as<int>(), M,
                       C, ys);
    MSR_CHECK_LAUNCH();
    BpTileArgs a{plan.items.as<int4>(), plan.pairs.as<int>(), plan.wts.as<float>(),
                 plan.cell.as<int2>(), plan.frame_s.as<int>(), ys, M, nch, g, grid, slab};
    MSR_REQUIRE(plan.nitems < ((int64_t)1 << 31), MSR_ERR_UNSUPPORTED, "too many work items");
    dim3 grd((unsigned)plan.nitems, (unsigned)ceil_div(nch, BP_CH));
    hipLaunchKernelGGL(bp_tile_kernel, grd, dim3(64), 0, st, a, src);
    MSR_CHECK_LAUNCH();
    if (plan.nsplit > 0) {
        hipLaunchKernelGGL(bp_split_reduce_kernel, dim3((unsigned)plan.nsplit, (unsigned)nch),
                           dim3(256), 0, st, plan.split_tiles.as<int2>(),
                           plan.split_nitems.as<int>(), plan.items.as<int4>(), slab, nch, g, grid);
        MSR_CHECK_LAUNCH();
    }
}

}  // namespace msr
