// Fused 2D Toeplitz subspace normal operator: passes A and C plus dispatch.
// Pass B (the kernel multiply) lives in apply2d_b.hip.
#include "apply2d.hpp"
#include "apply2d_b.hpp"
#include "fft.hpp"

namespace msr {

constexpr int RT2 = 16;  // rows per tile in passes A and C (16 * 8 B = 128 B segments)

template <int NX2>
constexpr int rowpass_block() {
    return (RT2 * FFTPlan<NX2>::T) < 256 ? RT2 * FFTPlan<NX2>::T : 256;
}

// ----------------------------------------------------------------- pass A
template <int NX2>
__global__ void __launch_bounds__(256)
pass_a_2d(const float2* __restrict__ b, const float2* __restrict__ cmaps,
          float2* __restrict__ T, const float2* __restrict__ tw, int nx, int ny, int K,
          const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int TP = FFTPlan<NX2>::T;
    constexpr int BLK = rowpass_block<NX2>();
    constexpr int F = BLK / TP;          // rows transformed concurrently
    constexpr int LD = lds_line<NX2>() + 1;  // LDS row pitch (odd: transposed access)
    __shared__ float2 s_tw[NX2];
    __shared__ float2 s_rows[RT2 * LD];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * RT2;
    const int k = blockIdx.y;
    const int c = blockIdx.z;
    load_twiddles<NX2>(s_tw, tw);
    const float2* bk = b + (size_t)k * nx * ny;
    const float2* cm = cmaps ? cmaps + (size_t)c * nx * ny : nullptr;
    // load RT2 rows, multiply by the coil map, zero-pad to NX2
    for (int e = tid; e < RT2 * NX2; e += BLK) {
        const int yl = e / NX2, xx = e - yl * NX2;
        const int y = y0 + yl;
        float2 v = make_float2(0.f, 0.f);
        if (xx < nx && y < ny) {
            v = bk[(size_t)y * nx + xx];
            if (cm) v = cmul(cm[(size_t)y * nx + xx], v);
        }
        s_rows[yl * LD + lds_pad(xx)] = v;
    }
    __syncthreads();
    const int g = tid / TP, t = tid - g * TP;
#pragma unroll
    for (int r0 = 0; r0 < RT2; r0 += F) fft_lds<NX2, -1>(s_rows + (r0 + g) * LD, s_tw, t);
    // transposed write: T[c][k][qx][y], y fastest
    float2* Tk = T + ((size_t)c * K + k) * NX2 * ny;
    for (int e = tid; e < RT2 * NX2; e += BLK) {
        const int yl = e % RT2, qx = e / RT2;
        const int y = y0 + yl;
        if (y < ny) Tk[(size_t)qx * ny + y] = s_rows[yl * LD + lds_pad(qx)];
    }
}

// ----------------------------------------------------------------- pass C
template <int NX2>
__global__ void __launch_bounds__(256)
pass_c_2d(const float2* __restrict__ T, const float2* __restrict__ cmaps,
          float2* __restrict__ partial, const float2* __restrict__ tw, int nx, int ny, int K,
          int C, int G, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int TP = FFTPlan<NX2>::T;
    constexpr int BLK = rowpass_block<NX2>();
    constexpr int F = BLK / TP;
    constexpr int LD = lds_line<NX2>() + 1;
    constexpr int NXC = NX2 / 2;                 // = nx
    constexpr int EA = (RT2 * NXC + BLK - 1) / BLK;
    __shared__ float2 s_tw[NX2];
    __shared__ float2 s_rows[RT2 * LD];
    const int tid = threadIdx.x;
    const int y0 = blockIdx.x * RT2;
    const int k = blockIdx.y;
    const int grp = blockIdx.z;
    const int c0 = (int)((int64_t)C * grp / G), c1 = (int)((int64_t)C * (grp + 1) / G);
    load_twiddles<NX2>(s_tw, tw);
    float2 acc[EA];
#pragma unroll
    for (int i = 0; i < EA; ++i) acc[i] = make_float2(0.f, 0.f);
    const int g = tid / TP, t = tid - g * TP;
    for (int c = c0; c < c1; ++c) {
        const float2* Tk = T + ((size_t)c * K + k) * NX2 * ny;
        for (int e = tid; e < RT2 * NX2; e += BLK) {
            const int yl = e % RT2, qx = e / RT2;
            const int y = y0 + yl;
            s_rows[yl * LD + lds_pad(qx)] = (y < ny) ? Tk[(size_t)qx * ny + y] : make_float2(0.f, 0.f);
        }
        __syncthreads();
#pragma unroll
        for (int r0 = 0; r0 < RT2; r0 += F) fft_lds<NX2, +1>(s_rows + (r0 + g) * LD, s_tw, t);
        const float2* cm = cmaps ? cmaps + (size_t)c * nx * ny : nullptr;
#pragma unroll
        for (int i = 0; i < EA; ++i) {
            const int e = tid + i * BLK;
            if (e < RT2 * NXC) {
                const int yl = e / NXC, xx = e - yl * NXC;
                const int y = y0 + yl;
                if (y < ny) {
                    float2 v = s_rows[yl * LD + lds_pad(xx)];
                    if (cm) cfmac(acc[i], cm[(size_t)y * nx + xx], v);
                    else acc[i] = cadd(acc[i], v);
                }
            }
        }
        __syncthreads();
    }
    float2* pk = partial + ((size_t)grp * K + k) * nx * ny;
#pragma unroll
    for (int i = 0; i < EA; ++i) {
        const int e = tid + i * BLK;
        if (e < RT2 * NXC) {
            const int yl = e / NXC, xx = e - yl * NXC;
            const int y = y0 + yl;
            if (y < ny) pk[(size_t)y * nx + xx] = acc[i];
        }
    }
}

// ----------------------------------------------------------------- dispatch
void launch_apply2d_passes(const Apply2DArgs& a, const float2* b, hipStream_t st,
                           const int* skip) {
    const int NX2 = 2 * a.nx, NY2 = 2 * a.ny;
    const int tiles = (a.ny + RT2 - 1) / RT2;
    bool ok = dispatch_fft_len(NX2, [&]<int N>() {
        constexpr int BLK = rowpass_block<N>();
        hipLaunchKernelGGL(pass_a_2d<N>, dim3(tiles, a.K, a.C), dim3(BLK), 0, st, b, a.cmaps,
                           a.T, a.tw_x, a.nx, a.ny, a.K, skip);
    });
    MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "unsupported grid size 2Nx = %d", NX2);
    MSR_CHECK_LAUNCH();
    launch_pass_b_2d(a, st, skip);
    ok = dispatch_fft_len(NX2, [&]<int N>() {
        constexpr int BLK = rowpass_block<N>();
        hipLaunchKernelGGL(pass_c_2d<N>, dim3(tiles, a.K, a.G), dim3(BLK), 0, st, a.T, a.cmaps,
                           a.partial, a.tw_x, a.nx, a.ny, a.K, a.C, a.G, skip);
    });
    MSR_CHECK_LAUNCH();
    (void)NY2;
}

}  // namespace msr
