// Tile-sorted LDS spreading with the ES kernel (see spread.hpp).
//
// Point coordinate on a grid of n cells: s = k * n in double (exact for the
// float32 trajectory), split into the wrapped floor cell c = mod(floor(s), n)
// and the fraction f = s - floor(s) in [0, 1).  The W = 8 kernel cells are the
// offsets o = -3..4 from c with weight phi(o - f) (zero outside |o - f| < 4),
// so the point's support stays within 4 cells of its tile: a tile of TB cells
// per dim is accumulated in a (TB + 8)^d LDS block.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "spread.hpp"

namespace msr {

template <int D> struct TileShape;
template <> struct TileShape<2> {
    static constexpr int X = 32, Y = 32, Z = 1;
    static constexpr int CHG = 8;   // channels per workgroup (one wavefront each), ComplexF32
};
#ifndef MSR_T3X
#define MSR_T3X 16
#define MSR_T3Y 8
#define MSR_T3Z 8
#define MSR_T3CHG 2
#endif
template <> struct TileShape<3> {
    static constexpr int X = MSR_T3X, Y = MSR_T3Y, Z = MSR_T3Z;
    static constexpr int CHG = MSR_T3CHG;
};
// channels per workgroup for the accumulation type AT (LDS: ComplexF64 tiles
// take twice the space)
template <int D, class AT>
constexpr int chg_of() { return sizeof(AT) == 16 ? TileShape<D>::CHG / 2 : TileShape<D>::CHG; }

template <class RT>
__device__ __forceinline__ RT es_phi(RT t) {
    const RT z = t * (RT)(2.0 / SPREAD_W);
    const RT a = (RT)1 - z * z;
    return a > (RT)0 ? exp((RT)SPREAD_BETA * (sqrt(a) - (RT)1)) : (RT)0;
}

template <class RT>
__device__ __forceinline__ void point_cell(float k, int64_t n, int64_t& cell, RT& frac) {
    const double s = (double)k * (double)n;
    const double fl = floor(s);
    frac = (RT)(s - fl);
    int64_t c = (int64_t)fl % n;
    if (c < 0) c += n;
    cell = c;
}

struct PlanGeom {
    int64_t n[3];
    int tb[3];
    int nb[3];
};

// ------------------------------------------------------------------ sort
__global__ void bin_count_kernel(const float* __restrict__ kc, int d, int64_t M, PlanGeom g,
                                 int* __restrict__ counts, int* __restrict__ binid) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < M;
         j += (int64_t)gridDim.x * blockDim.x) {
        int b = 0, mul = 1;
        for (int a = 0; a < d; ++a) {
            int64_t c;
            float f;
            point_cell<float>(kc[j * d + a], g.n[a], c, f);
            b += (int)(c / g.tb[a]) * mul;
            mul *= g.nb[a];
        }
        binid[j] = b;
        atomicAdd(&counts[b], 1);
    }
}

// exclusive scan of counts into offs[0..nb], single block
__global__ void bin_scan_kernel(const int* __restrict__ counts, int64_t nb, int* __restrict__ offs) {
    __shared__ int s[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int v0 = i < nb ? counts[i] : 0;
        s[threadIdx.x] = v0;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < nb) offs[i] = carry + s[threadIdx.x] - v0;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) offs[nb] = carry;
}

__global__ void izero_kernel(int* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

static void izero(int* p, int64_t n, hipStream_t st) {
    int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(izero_kernel, dim3(grid), dim3(256), 0, st, p, n);
    MSR_CHECK_LAUNCH();
}

__global__ void iota_kernel(int* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int)i;
}

// Points sorted by tile with a stable radix sort (ties keep sample order), so
// the spreading order, and with it the FP32 result, is reproducible.
const SpreadPlan& spread_plan(const msr_traj* traj, const int64_t* n, hipStream_t st) {
    std::lock_guard<std::recursive_mutex> lock(traj->mu);
    std::array<int64_t, 3> key{n[0], traj->d > 1 ? n[1] : 1, traj->d > 2 ? n[2] : 1};
    auto it = traj->plans.find(key);
    if (it != traj->plans.end()) return *it->second;
    auto p = std::make_unique<SpreadPlan>();
    p->d = traj->d;
    const int tb2[3] = {TileShape<2>::X, TileShape<2>::Y, 1};
    const int tb3[3] = {TileShape<3>::X, TileShape<3>::Y, TileShape<3>::Z};
    p->nbins = 1;
    for (int a = 0; a < 3; ++a) {
        p->n[a] = key[a];
        p->tb[a] = a < traj->d ? (traj->d == 2 ? tb2[a] : tb3[a]) : 1;
        p->nb[a] = (int)ceil_div(p->n[a], p->tb[a]);
        p->nbins *= p->nb[a];
    }
    MSR_REQUIRE(traj->M < (int64_t)INT32_MAX, MSR_ERR_UNSUPPORTED, "more than 2^31 samples");
    MSR_REQUIRE(p->nbins < (int64_t)INT32_MAX, MSR_ERR_UNSUPPORTED, "grid too large");
    PlanGeom g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = p->n[a];
        g.tb[a] = p->tb[a];
        g.nb[a] = p->nb[a];
    }
    const int M = (int)traj->M;
    DevBuf counts, binid, binsorted, idx, temp;
    counts.alloc(sizeof(int) * p->nbins);
    binid.alloc(sizeof(int) * M);
    binsorted.alloc(sizeof(int) * M);
    idx.alloc(sizeof(int) * M);
    p->perm.alloc(sizeof(int) * M);
    p->offs.alloc(sizeof(int) * (p->nbins + 1));
    izero(counts.as<int>(), p->nbins, st);
    int grid = (int)std::min<int64_t>((traj->M + 255) / 256, 8192);
    hipLaunchKernelGGL(bin_count_kernel, dim3(grid), dim3(256), 0, st, traj->kc.as<float>(),
                       traj->d, traj->M, g, counts.as<int>(), binid.as<int>());
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(256), 0, st, idx.as<int>(), traj->M);
    MSR_CHECK_LAUNCH();
    int bits = 1;
    while ((int64_t(1) << bits) < p->nbins) ++bits;
    size_t tbytes = 0;
    MSR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, binid.as<int>(),
                                               binsorted.as<int>(), idx.as<int>(),
                                               p->perm.as<int>(), M, 0, bits, st));
    temp.alloc(std::max<size_t>(tbytes, 16));
    MSR_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tbytes, binid.as<int>(),
                                               binsorted.as<int>(), idx.as<int>(),
                                               p->perm.as<int>(), M, 0, bits, st));
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, st, counts.as<int>(), p->nbins,
                       p->offs.as<int>());
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipStreamSynchronize(st));  // temporaries are freed on return
    auto& ref = *p;
    traj->plans[key] = std::move(p);
    return ref;
}

// ------------------------------------------------------------------ values
__device__ __forceinline__ float2 basis_value(const ValueSrc& s, int t, int k) {
    if (!s.U) return make_float2(t == k ? 1.f : 0.f, 0.f);
    if (s.u_real) return make_float2(reinterpret_cast<const float*>(s.U)[t + (int64_t)k * s.nt], 0.f);
    return reinterpret_cast<const float2*>(s.U)[t + (int64_t)k * s.nt];
}

// Value of channel ch at point j in the accumulation type AT (the product of
// two ComplexF32 basis values is exact in ComplexF64).
template <class AT>
__device__ __forceinline__ AT channel_value(const ValueSrc& s, const float* kc, int d,
                                            const int* frame, int64_t j, int ch) {
    const int t = frame[j];
    if (s.kind == 0) {
        const int c = ch / s.K, k = ch - c * s.K;
        const int64_t idx = s.sel ? s.sel[j] : j;
        float2 y = s.data[idx + (int64_t)c * s.nsel];
        float2 v = cmulc(basis_value(s, t, k), y);  // y * conj(U[t, k])
        if (s.dcf_mode == 1 || s.dcf_mode == 2) {
            float r2 = 0.f;
            for (int a = 0; a < d; ++a) r2 += kc[j * d + a] * kc[j * d + a];
            const float w = s.dcf_mode == 1 ? sqrtf(r2) : r2;
            v = cscale(v, w);
        } else if (s.dcf_mode == 3) {
            v = cscale(v, s.dcf[j]);
        }
        return cto<AT>(v);
    }
    const AT a = cto<AT>(basis_value(s, t, s.pair_j[ch]));
    const AT b = cto<AT>(basis_value(s, t, s.pair_k[ch]));
    return cmulc(a, b);  // conj(U[t, j]) * U[t, k]
}

// ------------------------------------------------------------------ spread
// 2D: one wavefront per value channel; 3D: eight per channel, wave w owning
// the padded tile's z planes with lz = w (mod 8), so that of the 8 planes a
// point touches each wave updates exactly one.  The waves of a workgroup walk
// the tile's points in sorted order, 64 points per chunk, with the chunk's
// kernel weights and channel values staged in LDS.  Lane (ox, oy) owns kernel
// cell (c + ox, c + oy) of the current point, so every LDS cell is updated by
// one wave in program order: the accumulation order is fixed and the result is
// bit-reproducible.
constexpr int CHUNK = 64;
template <int D> constexpr int waves_per_channel() { return D == 3 ? 8 : 1; }

// 2D, and 3D with a slab: every padded tile goes to slab[ch - chan0][bin][cell]
// with plain stores (tiles without points write nothing), summed afterwards by
// a gather over the covering tiles.  3D without a slab (slab == nullptr): the
// tile is added into the grid with atomics.
template <int D, class AT>
__global__ void __launch_bounds__(1024)
spread_tiles_kernel(const float* __restrict__ kc, const int* __restrict__ frame,
                    const int* __restrict__ perm, const int* __restrict__ offs, PlanGeom g,
                    ValueSrc src, int nch, int chan0, AT* __restrict__ slab,
                    AT* __restrict__ grid) {
    using TS = TileShape<D>;
    using RT = real_t<AT>;
    constexpr int H = SPREAD_HALO;
    constexpr int LX = TS::X + 2 * H, LY = TS::Y + 2 * H, LZ = D == 3 ? TS::Z + 2 * H : 1;
    constexpr int CELLS = LX * LY * LZ;
    constexpr int CHG = chg_of<D, AT>();
    __shared__ AT tile[CHG * CELLS];
    __shared__ RT s_w[CHUNK][D][8];
    __shared__ int s_base[CHUNK][D];
    __shared__ AT s_val[CHG][CHUNK];
    const int bin = blockIdx.x;
    const int ch0 = chan0 + blockIdx.y * CHG;
    if (D == 3 && slab && offs[bin] == offs[bin + 1]) return;  // empty tile: nothing to store
    const int bx = bin % g.nb[0];
    const int by = (bin / g.nb[0]) % g.nb[1];
    const int bz = D == 3 ? bin / (g.nb[0] * g.nb[1]) : 0;
    const int64_t t0[3] = {(int64_t)bx * TS::X, (int64_t)by * TS::Y, (int64_t)bz * TS::Z};
    for (int e = threadIdx.x; e < CHG * CELLS; e += blockDim.x) tile[e] = cz<AT>();

    const int lane = threadIdx.x & 63;
    const int wave = (threadIdx.x >> 6) / waves_per_channel<D>();  // local channel
    const int pz = (threadIdx.x >> 6) & 7;                          // 3D: owned plane residue
    const int ox = lane & 7, oy = lane >> 3;
    const int p0 = offs[bin], p1 = offs[bin + 1];
    for (int c0 = p0; c0 < p1; c0 += CHUNK) {
        const int np = min(CHUNK, p1 - c0);
        __syncthreads();
        // kernel weights phi(o - 3 - f) and tile-local base cells of the chunk
        for (int e = threadIdx.x; e < CHUNK * D * 8; e += blockDim.x) {
            const int pl = e / (D * 8), r = e - pl * (D * 8), a = r >> 3, o = r & 7;
            if (pl < np) {
                const int64_t j = perm[c0 + pl];
                int64_t cell;
                RT f;
                point_cell<RT>(kc[j * D + a], g.n[a], cell, f);
                s_w[pl][a][o] = es_phi<RT>((RT)(o - 3) - f);
                if (o == 0) s_base[pl][a] = (int)(cell - t0[a]) + H - 3;
            }
        }
        for (int e = threadIdx.x; e < CHUNK * CHG; e += blockDim.x) {
            const int ch = e / CHUNK, pl = e - ch * CHUNK;
            if (pl < np && ch0 + ch < nch)
                s_val[ch][pl] = channel_value<AT>(src, kc, D, frame, perm[c0 + pl], ch0 + ch);
        }
        __syncthreads();
        if (ch0 + wave < nch) {
            AT* tw = tile + wave * CELLS;
            for (int pl = 0; pl < np; ++pl) {
                const AT v = s_val[wave][pl];
                const RT wxy = s_w[pl][0][ox] * s_w[pl][1][oy];
                const int lx = s_base[pl][0] + ox, ly = s_base[pl][1] + oy;
                if (D == 2) {
                    AT* cell = tw + ly * LX + lx;
                    cell->x += wxy * v.x;
                    cell->y += wxy * v.y;
                } else {
                    const int lz0 = s_base[pl][D - 1];  // >= 0
                    const int oz = (pz - lz0) & 7;      // the offset landing on an owned plane
                    const RT w = wxy * s_w[pl][D - 1][oz];
                    AT* cell = tw + ((lz0 + oz) * LY + ly) * LX + lx;
                    cell->x += w * v.x;
                    cell->y += w * v.y;
                }
            }
        }
    }
    __syncthreads();
    const int chg = min(CHG, nch - ch0);
    if (D == 2 || slab) {
        // padded tile -> slab[ch][bin][cell] (plain stores; summed by gather)
        const int64_t nbins = (int64_t)g.nb[0] * g.nb[1] * (D == 3 ? g.nb[2] : 1);
        for (int e = threadIdx.x; e < chg * CELLS; e += blockDim.x) {
            const int ch = e / CELLS, r = e - ch * CELLS;
            slab[((int64_t)(ch0 - chan0 + ch) * nbins + bin) * CELLS + r] = tile[e];
        }
    } else {
        const int64_t G = g.n[0] * g.n[1] * g.n[2];
        for (int e = threadIdx.x; e < chg * CELLS; e += blockDim.x) {
            const AT v = tile[e];
            if (v.x == (RT)0 && v.y == (RT)0) continue;
            const int ch = e / CELLS;
            int r = e - ch * CELLS;
            const int lx = r % LX;
            r /= LX;
            const int ly = r % LY;
            const int lz = r / LY;
            int64_t gx = (t0[0] - H + lx) % g.n[0];
            if (gx < 0) gx += g.n[0];
            int64_t gy = (t0[1] - H + ly) % g.n[1];
            if (gy < 0) gy += g.n[1];
            int64_t gz = (t0[2] - H + lz) % g.n[2];
            if (gz < 0) gz += g.n[2];
            AT* dst = grid + (int64_t)(ch0 + ch) * G + (gz * g.n[1] + gy) * g.n[0] + gx;
            atomicAdd(&dst->x, v.x);
            atomicAdd(&dst->y, v.y);
        }
    }
}

// (tile, local) pairs of one dimension whose padded range covers grid cell gc,
// in increasing tile order (fixed summation order)
__device__ __forceinline__ int covering(int64_t gc, int64_t n, int tb, int nb, int* tl, int* loc) {
    constexpr int H = SPREAD_HALO;
    const int span = tb + 2 * H;
    int cnt = 0;
    const int b0 = (int)(gc / tb);
    const int nc = nb <= 3 ? nb : 3;
    for (int q = 0; q < nc; ++q) {
        int b = nb <= 3 ? q : b0 - 1 + q;
        b = (b % nb + nb) % nb;
        int64_t l = (gc - ((int64_t)b * tb - H)) % n;
        if (l < 0) l += n;
        for (; l < span && cnt < 6; l += n) {
            tl[cnt] = b;
            loc[cnt] = (int)l;
            ++cnt;
        }
    }
    // insertion sort by tile (then local) so the order is independent of b0
    for (int i = 1; i < cnt; ++i)
        for (int k = i; k > 0 && (tl[k - 1] > tl[k] || (tl[k - 1] == tl[k] && loc[k - 1] > loc[k])); --k) {
            int t = tl[k]; tl[k] = tl[k - 1]; tl[k - 1] = t;
            t = loc[k]; loc[k] = loc[k - 1]; loc[k - 1] = t;
        }
    return cnt;
}

template <class AT>
__global__ void gather_tiles_2d(const AT* __restrict__ slab, PlanGeom g, int nch,
                                AT* __restrict__ grid) {
    constexpr int H = SPREAD_HALO;
    constexpr int LX = TileShape<2>::X + 2 * H, LY = TileShape<2>::Y + 2 * H;
    constexpr int CELLS = LX * LY;
    const int64_t G = g.n[0] * g.n[1];
    const int64_t nbins = (int64_t)g.nb[0] * g.nb[1];
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < G * nch;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(e / G);
        const int64_t q = e - (int64_t)ch * G;
        const int64_t gy = q / g.n[0], gx = q - gy * g.n[0];
        int tx[6], lxs[6], ty[6], lys[6];
        const int nx = covering(gx, g.n[0], TileShape<2>::X, g.nb[0], tx, lxs);
        const int ny = covering(gy, g.n[1], TileShape<2>::Y, g.nb[1], ty, lys);
        AT acc = cz<AT>();
        const AT* sc = slab + (int64_t)ch * nbins * CELLS;
        for (int b = 0; b < ny; ++b)
            for (int a = 0; a < nx; ++a) {
                const AT v = sc[((int64_t)ty[b] * g.nb[0] + tx[a]) * CELLS + lys[b] * LX + lxs[a]];
                acc.x += v.x;
                acc.y += v.y;
            }
        grid[e] = acc;
    }
}

// grid[chan0 + ch][cell] = sum over the non-empty tiles whose padded range
// covers the cell (fixed order: z, y, x tile order), for nc channels
template <class AT>
__global__ void gather_tiles_3d(const AT* __restrict__ slab, const int* __restrict__ offs,
                                PlanGeom g, int nc, int chan0, AT* __restrict__ grid) {
    constexpr int H = SPREAD_HALO;
    using TS = TileShape<3>;
    constexpr int LX = TS::X + 2 * H, LY = TS::Y + 2 * H, LZ = TS::Z + 2 * H;
    constexpr int CELLS = LX * LY * LZ;
    const int64_t G = g.n[0] * g.n[1] * g.n[2];
    const int64_t nbins = (int64_t)g.nb[0] * g.nb[1] * g.nb[2];
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < G * nc;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(e / G);
        const int64_t q = e - (int64_t)ch * G;
        const int64_t gx = q % g.n[0];
        const int64_t r = q / g.n[0];
        const int64_t gy = r % g.n[1], gz = r / g.n[1];
        int tx[6], lxs[6], ty[6], lys[6], tz[6], lzs[6];
        const int nx = covering(gx, g.n[0], TS::X, g.nb[0], tx, lxs);
        const int ny = covering(gy, g.n[1], TS::Y, g.nb[1], ty, lys);
        const int nz = covering(gz, g.n[2], TS::Z, g.nb[2], tz, lzs);
        AT acc = cz<AT>();
        const AT* sc = slab + (int64_t)ch * nbins * CELLS;
        for (int c = 0; c < nz; ++c)
            for (int b = 0; b < ny; ++b)
                for (int a = 0; a < nx; ++a) {
                    const int64_t bin = ((int64_t)tz[c] * g.nb[1] + ty[b]) * g.nb[0] + tx[a];
                    if (offs[bin] == offs[bin + 1]) continue;  // empty tile: not stored
                    const AT v = sc[bin * CELLS + ((int64_t)lzs[c] * LY + lys[b]) * LX + lxs[a]];
                    acc.x += v.x;
                    acc.y += v.y;
                }
        grid[(int64_t)(chan0 + ch) * G + q] = acc;
    }
}

// device bytes of the 3D spreading slab per pass (channels per pass chosen
// so the slab stays within this budget; above it for one channel group the
// tiles are added with atomics).  MSR_SPREAD3D_SLAB_BUDGET (bytes) overrides
// the 8 GiB default.
static size_t spread3d_slab_budget() {
    const char* e = getenv("MSR_SPREAD3D_SLAB_BUDGET");
    return e ? (size_t)strtoull(e, nullptr, 10) : (size_t)8 << 30;
}

size_t spread_workspace_bytes(const SpreadPlan& plan, int nch, bool f64) {
    if (plan.d != 2) return 0;
    constexpr int CELLS = (TileShape<2>::X + 2 * SPREAD_HALO) * (TileShape<2>::Y + 2 * SPREAD_HALO);
    return (f64 ? sizeof(double2) : sizeof(float2)) * (size_t)nch * plan.nbins * CELLS;
}

void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

template <class AT>
static void spread_t(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
                     AT* grid, AT* slab, hipStream_t st) {
    const int64_t G = plan.n[0] * plan.n[1] * plan.n[2];
    PlanGeom g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = plan.n[a];
        g.tb[a] = plan.tb[a];
        g.nb[a] = plan.nb[a];
    }
    if (traj->d == 2) {
        MSR_REQUIRE(slab != nullptr, MSR_ERR_ERROR, "2D spreading needs a workspace");
        constexpr int CHG = chg_of<2, AT>();
        dim3 grd((unsigned)plan.nbins, (unsigned)ceil_div(nch, CHG));
        hipLaunchKernelGGL((spread_tiles_kernel<2, AT>), grd, dim3(64 * CHG), 0, st,
                           traj->kc.as<float>(), traj->frame.as<int>(), plan.perm.as<int>(),
                           plan.offs.as<int>(), g, src, nch, 0, slab, grid);
        MSR_CHECK_LAUNCH();
        int gr = (int)std::min<int64_t>((G * nch + 255) / 256, 65536);
        hipLaunchKernelGGL(gather_tiles_2d<AT>, dim3(gr), dim3(256), 0, st, slab, g, nch, grid);
    } else {
        constexpr int CHG = chg_of<3, AT>();
        constexpr int CELLS = (TileShape<3>::X + 2 * SPREAD_HALO) *
                              (TileShape<3>::Y + 2 * SPREAD_HALO) * (TileShape<3>::Z + 2 * SPREAD_HALO);
        const size_t per_ch = sizeof(AT) * (size_t)plan.nbins * CELLS;
        // channels per pass: a multiple of the workgroup's channel count
        int cpp = (int)std::min<size_t>((size_t)nch, spread3d_slab_budget() / per_ch);
        cpp = cpp >= CHG ? cpp / CHG * CHG : cpp;
        if (cpp < CHG) {  // slab of one channel group above the budget: atomics
            fill_floats(grid, (int64_t)(sizeof(AT) / sizeof(float)) * G * nch, 0.f, st);
            dim3 grd((unsigned)plan.nbins, (unsigned)ceil_div(nch, CHG));
            hipLaunchKernelGGL((spread_tiles_kernel<3, AT>), grd, dim3(64 * 8 * CHG), 0, st,
                               traj->kc.as<float>(), traj->frame.as<int>(), plan.perm.as<int>(),
                               plan.offs.as<int>(), g, src, nch, 0, (AT*)nullptr, grid);
        } else {
            DevBuf work;
            work.alloc(per_ch * cpp);
            for (int c0 = 0; c0 < nch; c0 += cpp) {
                const int nc = std::min(cpp, nch - c0);
                dim3 grd((unsigned)plan.nbins, (unsigned)ceil_div(nc, CHG));
                hipLaunchKernelGGL((spread_tiles_kernel<3, AT>), grd, dim3(64 * 8 * CHG), 0, st,
                                   traj->kc.as<float>(), traj->frame.as<int>(),
                                   plan.perm.as<int>(), plan.offs.as<int>(), g, src,
                                   c0 + nc, c0, work.as<AT>(), grid);
                MSR_CHECK_LAUNCH();
                const int gr = (int)std::min<int64_t>((G * nc + 255) / 256, 65536);
                hipLaunchKernelGGL(gather_tiles_3d<AT>, dim3(gr), dim3(256), 0, st, work.as<AT>(),
                                   plan.offs.as<int>(), g, nc, c0, grid);
                MSR_CHECK_LAUNCH();
            }
            MSR_HIP(hipStreamSynchronize(st));  // the slab is released on return
        }
    }
    MSR_CHECK_LAUNCH();
}

void spread(const msr_traj* traj, const SpreadPlan& plan, const ValueSrc& src, int nch,
            void* grid, void* slab, bool f64, hipStream_t st) {
    if (f64) spread_t<double2>(traj, plan, src, nch, (double2*)grid, (double2*)slab, st);
    else spread_t<float2>(traj, plan, src, nch, (float2*)grid, (float2*)slab, st);
}

// ------------------------------------------------------------------ deapodisation
// 1/phihat(r), phihat(r) = int phi(t) cos(2 pi r t / n_os) dt (midpoint rule, FP64)
template <class OT>
__global__ void deapod_kernel(int n_os, int n_out, int sel, OT* __restrict__ out) {
    const int j = blockIdx.x;
    const int r = sel == 1 ? j - n_out / 2 : (j < (n_out + 1) / 2 ? j : j - n_out);
    constexpr int NQ = 8192;
    const double h = (double)SPREAD_W / NQ;
    double acc = 0.0;
    for (int m = threadIdx.x; m < NQ; m += blockDim.x) {
        const double t = -0.5 * SPREAD_W + (m + 0.5) * h;
        const double z = 2.0 * t / SPREAD_W;
        const double a = 1.0 - z * z;
        const double ph = a > 0.0 ? exp((double)SPREAD_BETA * (sqrt(a) - 1.0)) : 0.0;
        acc += ph * cos(2.0 * M_PI * (double)r * t / (double)n_os);
    }
    __shared__ double s[256];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = (OT)(1.0 / (s[0] * h));
}

template <class OT>
static const OT* deapod_table_t(msr_ctx* ctx, std::map<std::tuple<int, int, int>, OT*>& cache,
                                int n_os, int n_out, int sel) {
    std::lock_guard<std::mutex> guard(ctx->lock);
    auto key = std::make_tuple(n_os, n_out, sel);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    OT* d = nullptr;
    MSR_HIP(hipMalloc(&d, sizeof(OT) * n_out));
    hipLaunchKernelGGL(deapod_kernel<OT>, dim3(n_out), dim3(256), 0, nullptr, n_os, n_out, sel, d);
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipDeviceSynchronize());
    cache[key] = d;
    return d;
}

const float* deapod_table(msr_ctx* ctx, int n_os, int n_out, int sel) {
    return deapod_table_t<float>(ctx, ctx->deapod, n_os, n_out, sel);
}

const double* deapod_table64(msr_ctx* ctx, int n_os, int n_out, int sel) {
    return deapod_table_t<double>(ctx, ctx->deapod64, n_os, n_out, sel);
}

}  // namespace msr
