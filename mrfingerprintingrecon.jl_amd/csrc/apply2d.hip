// Fused 2D Toeplitz subspace normal operator: passes A and C plus dispatch.
// Pass B (the kernel multiply) lives in apply2d_b.hpp.  Every pass is
// templated on the complex arithmetic type CT (float2 or double2).
#include "apply2d.hpp"
#include "apply2d_b.hpp"
#include "fft.hpp"

namespace msr {

template <int NX2, class CT, class ST>
constexpr int rowpass_block() {
    return (rows_per_tile<CT, ST>() * FFTPlan<NX2>::T) < 256
               ? rows_per_tile<CT, ST>() * FFTPlan<NX2>::T
               : 256;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b % 8),
// each XCD with its own L2.  Blocks b and b + 8 get consecutive logical
// indices (when the block count divides by 8), so that consecutive logical
// blocks run on one XCD close in time.  Two logical orders per pass, the same
// blocks and arithmetic in either (the result is bit-identical): pass A
// (tile, k, c) with the tile fastest (neighbouring tiles of one plane written
// together) or (c, k, tile) with the coil fastest (a coefficient's row tile of
// b read once from HBM for all coils, a coil map tile once for all
// coefficients); pass C (tile, k, group) with the tile fastest or (k, group,
// tile) with the coefficient fastest (a group's coil map tiles read once for
// all coefficients).  Which is faster depends on the storage type and the
// shape (at 256^2 / 20 coils / K = 5 on MI355X: pass A 74 against 93 us
// tile-fastest with 16-byte storage, 66 against 70 us coil-fastest with 8-byte
// storage; pass C 45 against 64 us coefficient-fastest in ComplexF32); the
// operator's knobs carry the choice, measured by the tuner.
__device__ __forceinline__ int64_t xcd_order(int64_t lb, int64_t nblk) {
    if (nblk % 8 != 0) return lb;
    return (lb % 8) * (nblk / 8) + lb / 8;
}

// Passes A and C keep the FFT's first (A) or last (C) stage in registers when
// the line's plan is exact with an even first / last radix: the zero-padded
// half of a row is then the constant zero of half the first stage's inputs,
// and the cropped half of the inverse transform half the last stage's outputs.
// For 8-byte storage of the intermediate only: at 256^2 / 20 coils / K = 5 on
// MI355X the apply went from 0.191 to 0.184 ms (ComplexF32) and 0.272 to
// 0.266 ms (compact), but from 0.303 to 0.309 ms with ComplexF64 storage.
template <int N, class ST>
constexpr bool rowreg_first() {
    return sizeof(ST) <= 8 && fft_plan_exact<N>() && FFTPlan<N>::R[0] % 2 == 0;
}
template <int N, class ST>
constexpr bool rowreg_last() {
    return sizeof(ST) <= 8 && fft_plan_exact<N>() &&
           FFTPlan<N>::R[FFTPlan<N>::NST - 1] % 2 == 0;
}

// ----------------------------------------------------------------- pass A
// block lb = (tile, k, c) or (c, k, tile) (a.order_a); CT arithmetic, ST
// storage of T
template <int NX2, class CT, class ST>
__device__ __forceinline__ void pass_a_body(const Apply2DArgs& a, int64_t lb) {
    if (a.skip && *a.skip) return;
    constexpr int RT = rows_per_tile<CT, ST>();
    constexpr int TP = FFTPlan<NX2>::T;
    constexpr int BLK = rowpass_block<NX2, CT, ST>();
    constexpr int F = BLK / TP;              // rows transformed concurrently
    constexpr int LD = lds_line<NX2, CT>() + 1;  // LDS row pitch (odd: transposed access)
    // ComplexF64 twiddles are read from the global table (L1 / L2), which frees
    // LDS for one more workgroup per CU; ComplexF32 ones from an LDS copy
    __shared__ CT s_tw_lds[sizeof(CT) == 16 ? 1 : NX2];
    const CT* s_tw = sizeof(CT) == 16 ? static_cast<const CT*>(a.tw_x) : s_tw_lds;
    __shared__ CT s_rows[RT * LD];
    const int nx = a.nx, ny = a.ny, K = a.K;
    const int tiles = tiles2d(ny, RT);
    lb = xcd_order(lb, (int64_t)tiles * K * a.C);
    int c, k, tile;
    if (a.order_a == 0) {
        tile = (int)(lb % tiles);
        k = (int)((lb / tiles) % K);
        c = (int)(lb / ((int64_t)tiles * K));
    } else {
        c = (int)(lb % a.C);
        k = (int)((lb / a.C) % K);
        tile = (int)(lb / ((int64_t)a.C * K));
    }
    const int tid = threadIdx.x;
    const int y0 = tile * RT;
    if constexpr (sizeof(CT) != 16)
        load_twiddles<NX2>(s_tw_lds, static_cast<const CT*>(a.tw_x));
    const CT* bk = static_cast<const CT*>(a.b) + (size_t)k * nx * ny;
    const float2* cm = a.cmaps ? a.cmaps + (size_t)c * nx * ny : nullptr;
    const int g = tid / TP, t = tid - g * TP;
    if constexpr (rowreg_first<NX2, ST>()) {
        // each line's threads load their row (times the coil map) straight
        // into the first stage's layout, x = t + p TP + r NX2/R0 at reg[p R0 + r];
        // x < nx = NX2/2 exactly for r < R0/2, the rest is the zero padding
        constexpr int R0 = FFTPlan<NX2>::R[0], NB0 = NX2 / R0, NR = NX2 / TP;
#pragma unroll
        for (int r0 = 0; r0 < RT; r0 += F) {
            const int yl = r0 + g, y = y0 + yl;
            CT reg[NR];
#pragma unroll
            for (int p = 0; p < NB0 / TP; ++p)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    CT v = cz<CT>();
                    if (r < R0 / 2 && y < ny) {
                        const size_t o = (size_t)y * nx + t + p * TP + r * NB0;
                        v = bk[o];
                        if (cm) v = cmul(cto<CT>(cm[o]), v);
                    }
                    reg[p * R0 + r] = v;
                }
            fft_wave<NX2, -1, true, false>(s_rows + yl * LD, s_tw, t, true, reg);
        }
        __syncthreads();  // every row in LDS before the transposed write
    } else {
        // load RT rows, multiply by the coil map, zero-pad to NX2
        for (int e = tid; e < RT * NX2; e += BLK) {
            const int yl = e / NX2, xx = e - yl * NX2;
            const int y = y0 + yl;
            CT v = cz<CT>();
            if (xx < nx && y < ny) {
                v = bk[(size_t)y * nx + xx];
                if (cm) v = cmul(cto<CT>(cm[(size_t)y * nx + xx]), v);
            }
            s_rows[yl * LD + lds_pad<NX2, CT>(xx)] = v;
        }
        __syncthreads();
#pragma unroll
        for (int r0 = 0; r0 < RT; r0 += F) fft_lds<NX2, -1>(s_rows + (r0 + g) * LD, s_tw, t);
    }
    // transposed write: T[c][k][qx][y], y fastest
    ST* Tk = static_cast<ST*>(a.T) + ((size_t)c * K + k) * NX2 * ny;
    for (int e = tid; e < RT * NX2; e += BLK) {
        const int yl = e % RT, qx = e / RT;
        const int y = y0 + yl;
        if (y < ny) st_store<CT, ST>(Tk + (size_t)qx * ny + y, s_rows[yl * LD + lds_pad<NX2, CT>(qx)]);
    }
}

template <int NX2, class CT, class ST>
__global__ void __launch_bounds__(256) pass_a_2d(Apply2DArgs a) {
    pass_a_body<NX2, CT, ST>(a, blockIdx.x);
}

template <int NX2, class CT, class ST>
__global__ void __launch_bounds__(256)
pass_a_2d_batch(const Apply2DArgs* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    pass_a_body<NX2, CT, ST>(jobs[j], (int64_t)blockIdx.x - pre[j]);
}

// ----------------------------------------------------------------- pass C
// block lb = (tile, k, coil group) or (k, coil group, tile) (a.order_c)
template <int NX2, class CT, class ST>
__device__ __forceinline__ void pass_c_body(const Apply2DArgs& a, int64_t lb) {
    if (a.skip && *a.skip) return;
    constexpr int RT = rows_per_tile<CT, ST>();
    constexpr int TP = FFTPlan<NX2>::T;
    constexpr int BLK = rowpass_block<NX2, CT, ST>();
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NX2, CT>() + 1;
    constexpr int NXC = NX2 / 2;                 // = nx
    constexpr bool LASTREG = rowreg_last<NX2, ST>();
    constexpr int RL = FFTPlan<NX2>::R[FFTPlan<NX2>::NST - 1], NBL = NX2 / RL;
    constexpr int HALF = NX2 / TP / 2;           // cropped outputs per line and thread
    // accumulators: per thread EA elements of the tile (LDS crop), or HALF per
    // line it transforms (register crop)
    constexpr int EA = LASTREG ? (RT / F) * HALF : (RT * NXC + BLK - 1) / BLK;
    // ComplexF64 twiddles are read from the global table (L1 / L2), which frees
    // LDS for one more workgroup per CU; ComplexF32 ones from an LDS copy
    __shared__ CT s_tw_lds[sizeof(CT) == 16 ? 1 : NX2];
    const CT* s_tw = sizeof(CT) == 16 ? static_cast<const CT*>(a.tw_x) : s_tw_lds;
    __shared__ CT s_rows[RT * LD];
    const int nx = a.nx, ny = a.ny, K = a.K, C = a.C, G = a.G;
    const int tiles = tiles2d(ny, RT);
    lb = xcd_order(lb, (int64_t)tiles * K * G);
    int k, grp, tile;
    if (a.order_c == 0) {
        tile = (int)(lb % tiles);
        k = (int)((lb / tiles) % K);
        grp = (int)(lb / ((int64_t)tiles * K));
    } else {
        k = (int)(lb % K);
        grp = (int)((lb / K) % G);
        tile = (int)(lb / ((int64_t)K * G));
    }
    const int tid = threadIdx.x;
    const int y0 = tile * RT;
    const int c0 = (int)((int64_t)C * grp / G), c1 = (int)((int64_t)C * (grp + 1) / G);
    if constexpr (sizeof(CT) != 16)
        load_twiddles<NX2>(s_tw_lds, static_cast<const CT*>(a.tw_x));
    CT acc[EA];
#pragma unroll
    for (int i = 0; i < EA; ++i) acc[i] = cz<CT>();
    const int g = tid / TP, t = tid - g * TP;
    // the tile's rows of one coil, NLE elements per thread, prefetched into
    // registers one coil ahead
    constexpr int NLE = (RT * NX2 + BLK - 1) / BLK;
    ST pre[NLE];  // kept in the storage type (half the registers for cx64)
    auto prefetch = [&](int c) {
        const ST* Tk = static_cast<const ST*>(a.T) + ((size_t)c * K + k) * NX2 * ny;
#pragma unroll
        for (int r = 0; r < NLE; ++r) {
            const int e = tid + r * BLK;
            const int yl = e % RT, qx = e / RT;
            const int y = y0 + yl;
            if (e < RT * NX2 && y < ny) pre[r] = Tk[(size_t)qx * ny + y];
            else pre[r] = ST{};
        }
    };
    if (c0 < c1) prefetch(c0);
    for (int c = c0; c < c1; ++c) {
#pragma unroll
        for (int r = 0; r < NLE; ++r) {
            const int e = tid + r * BLK;
            if (e < RT * NX2) s_rows[(e % RT) * LD + lds_pad<NX2, CT>(e / RT)] = st_conv<CT, ST>(pre[r]);
        }
        __syncthreads();
        if (c + 1 < c1) prefetch(c + 1);  // in flight during this coil's transforms
        const float2* cm = a.cmaps ? a.cmaps + (size_t)c * nx * ny : nullptr;
        if constexpr (LASTREG) {
            // each line's last stage leaves X[t + p TP + r NX2/RL] in reg[p RL + r];
            // the crop keeps x < nx, i.e. r < RL/2; acc[(r0/F) HALF + p RL/2 + r]
#pragma unroll
            for (int r0 = 0; r0 < RT; r0 += F) {
                const int yl = r0 + g, y = y0 + yl;
                CT reg[NX2 / TP];
                fft_wave<NX2, +1, false, true>(s_rows + yl * LD, s_tw, t, true, reg);
#pragma unroll
                for (int p = 0; p < NBL / TP; ++p)
#pragma unroll
                    for (int r = 0; r < RL / 2; ++r) {
                        const int xx = t + p * TP + r * NBL;
                        CT& ac = acc[(r0 / F) * HALF + p * (RL / 2) + r];
                        if (y < ny) {
                            if (cm) cfmac(ac, cto<CT>(cm[(size_t)y * nx + xx]), reg[p * RL + r]);
                            else ac = cadd(ac, reg[p * RL + r]);
                        }
                    }
            }
        } else {
#pragma unroll
            for (int r0 = 0; r0 < RT; r0 += F) fft_lds<NX2, +1>(s_rows + (r0 + g) * LD, s_tw, t);
#pragma unroll
            for (int i = 0; i < EA; ++i) {
                const int e = tid + i * BLK;
                if (e < RT * NXC) {
                    const int yl = e / NXC, xx = e - yl * NXC;
                    const int y = y0 + yl;
                    if (y < ny) {
                        CT v = s_rows[yl * LD + lds_pad<NX2, CT>(xx)];
                        if (cm) cfmac(acc[i], cto<CT>(cm[(size_t)y * nx + xx]), v);
                        else acc[i] = cadd(acc[i], v);
                    }
                }
            }
        }
        __syncthreads();
    }
    CT* pk = static_cast<CT*>(a.partial) + ((size_t)grp * K + k) * nx * ny;
    if constexpr (LASTREG) {
#pragma unroll
        for (int r0 = 0; r0 < RT; r0 += F) {
            const int y = y0 + r0 + g;
#pragma unroll
            for (int p = 0; p < NBL / TP; ++p)
#pragma unroll
                for (int r = 0; r < RL / 2; ++r)
                    if (y < ny)
                        pk[(size_t)y * nx + t + p * TP + r * NBL] =
                            acc[(r0 / F) * HALF + p * (RL / 2) + r];
        }
    } else {
#pragma unroll
        for (int i = 0; i < EA; ++i) {
            const int e = tid + i * BLK;
            if (e < RT * NXC) {
                const int yl = e / NXC, xx = e - yl * NXC;
                const int y = y0 + yl;
                if (y < ny) pk[(size_t)y * nx + xx] = acc[i];
            }
        }
    }
}

template <int NX2, class CT, class ST>
__global__ void __launch_bounds__(256) pass_c_2d(Apply2DArgs a) {
    pass_c_body<NX2, CT, ST>(a, blockIdx.x);
}

template <int NX2, class CT, class ST>
__global__ void __launch_bounds__(256)
pass_c_2d_batch(const Apply2DArgs* __restrict__ jobs, const int* __restrict__ pre, int nj) {
    const int j = batch_job(pre, nj, blockIdx.x);
    pass_c_body<NX2, CT, ST>(jobs[j], (int64_t)blockIdx.x - pre[j]);
}

// ----------------------------------------------------------------- dispatch
static unsigned grid1d(int64_t nblocks) {
    MSR_REQUIRE(nblocks >= 1 && nblocks < (int64_t)1 << 31, MSR_ERR_ARGUMENT,
                "launch of %lld blocks", (long long)nblocks);
    return (unsigned)nblocks;
}

// the row passes need 2Nx <= 512 (LDS); pass B is limited the same way
template <int N>
constexpr bool row_len_ok() { return N <= 512; }

// Calls f.template operator()<CT, ST>() for the arithmetic / storage pair of
// the apply mode.
template <class F>
static void dispatch_mode(int f64, int compact, F&& f) {
    if (!f64) f.template operator()<float2, float2>();
    else if (compact) f.template operator()<double2, cx64>();
    else f.template operator()<double2, double2>();
}

static void launch_a(const Apply2DArgs& a, hipStream_t st) {
    const int NX2 = 2 * a.nx;
    bool ok = dispatch_fft_len(NX2, [&]<int N>() {
        if constexpr (row_len_ok<N>())
            dispatch_mode(a.f64, a.compact, [&]<class CT, class ST>() {
                hipLaunchKernelGGL((pass_a_2d<N, CT, ST>), dim3(grid1d(blocks_a2d(a))),
                                   dim3(rowpass_block<N, CT, ST>()), 0, st, a);
            });
        else
            fail(MSR_ERR_UNSUPPORTED, "Toeplitz grid 2Nx = %d above 512", N);
    });
    MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Nx = %d", NX2);
    MSR_CHECK_LAUNCH();
}

static void launch_c(const Apply2DArgs& a, hipStream_t st) {
    dispatch_fft_len(2 * a.nx, [&]<int N>() {
        if constexpr (row_len_ok<N>())
            dispatch_mode(a.f64, a.compact, [&]<class CT, class ST>() {
                hipLaunchKernelGGL((pass_c_2d<N, CT, ST>), dim3(grid1d(blocks_c2d(a))),
                                   dim3(rowpass_block<N, CT, ST>()), 0, st, a);
            });
    });
    MSR_CHECK_LAUNCH();
}

void launch_apply2d_passes(const Apply2DArgs& a, hipStream_t st) {
    launch_a(a, st);
    launch_pass_b_2d(a, st);
    launch_c(a, st);
}

int pass_b_colblocks(int nx2, int ny2, int K, int f64) {
    int r = 0;
#define MSR_CALLB(KK) r = pass_b_2d_colblocks_k##KK(nx2, ny2, f64)
    MSR_PASSB_SWITCH(K, MSR_CALLB)
#undef MSR_CALLB
    return r;
}

void launch_pass_b_2d_batch(int ny2, int K, int lam_real, int am, const Apply2DArgs* jobs,
                            const int* pre, int nj, int64_t nblocks, hipStream_t st) {
    // the template arguments come from (ny2, K, lam_real, am); the job table
    // itself stays on the device
    Apply2DArgs shape{};
    shape.ny = ny2 / 2;
    shape.K = K;
    shape.lam_real = lam_real;
    shape.f64 = am != AM_F32;
    shape.compact = am == AM_F64C;
#define MSR_CALLB(KK) pass_b_2d_k##KK(shape, jobs, pre, nj, nblocks, st)
    MSR_PASSB_SWITCH(K, MSR_CALLB)
#undef MSR_CALLB
    MSR_CHECK_LAUNCH();
}

void launch_pass_a_2d_batch(int nx2, int am, const Apply2DArgs* jobs, const int* pre, int nj,
                            int64_t nblocks, hipStream_t st) {
    bool ok = dispatch_fft_len(nx2, [&]<int N>() {
        if constexpr (row_len_ok<N>())
            dispatch_mode(am != AM_F32, am == AM_F64C, [&]<class CT, class ST>() {
                hipLaunchKernelGGL((pass_a_2d_batch<N, CT, ST>), dim3(grid1d(nblocks)),
                                   dim3(rowpass_block<N, CT, ST>()), 0, st, jobs, pre, nj);
            });
    });
    MSR_REQUIRE(ok && nx2 <= 512, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Nx = %d", nx2);
    MSR_CHECK_LAUNCH();
}

void launch_pass_c_2d_batch(int nx2, int am, const Apply2DArgs* jobs, const int* pre, int nj,
                            int64_t nblocks, hipStream_t st) {
    bool ok = dispatch_fft_len(nx2, [&]<int N>() {
        if constexpr (row_len_ok<N>())
            dispatch_mode(am != AM_F32, am == AM_F64C, [&]<class CT, class ST>() {
                hipLaunchKernelGGL((pass_c_2d_batch<N, CT, ST>), dim3(grid1d(nblocks)),
                                   dim3(rowpass_block<N, CT, ST>()), 0, st, jobs, pre, nj);
            });
    });
    MSR_REQUIRE(ok && nx2 <= 512, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Nx = %d", nx2);
    MSR_CHECK_LAUNCH();
}

}  // namespace msr
