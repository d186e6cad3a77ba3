// Forward subspace model (type-2 NUFFT side), the data simulation of the
// reference tests (test/reconstruct_radial.jl:62-74: for every coil and frame,
// NonuniformFFTs exec_type2! of S_c * sum_k U[t, k] x_k):
//   y_{j,c} = sum_k U[t_j, k] sum_n S_c[n] x_k[n] exp(-2 pi i k_j . (n - N/2))
// On the device: z_{c,k} = S_c x_k; centred placement on the sigma = 2 grid with
// deapodisation 1/phihat fused into the first line pass; forward FFTs; then ES
// interpolation (W = 8) at the samples, where each lane of a wavefront owns one
// kernel cell, folds the K channels with U[t_j, k] in registers, and the
// wavefront reduces the 64 (8x8, z looped in 3D) cell contributions.
#include <algorithm>

#include "fft.hpp"
#include "fftlines.hpp"
#include "spread.hpp"

namespace msr {

void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

// Z[(c*K + k)][v] = S_c[v] x_k[v]
__global__ void expand_kernel(const float2* __restrict__ x, const float2* __restrict__ cmaps, int C,
                              int K, int64_t V, float2* __restrict__ Z) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)C * K * V;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = e % V;
        const int64_t ck = e / V;
        const int k = (int)(ck % K), c = (int)(ck / K);
        const float2 xv = x[(int64_t)k * V + v];
        Z[e] = cmaps ? cmul(cmaps[(int64_t)c * V + v], xv) : xv;
    }
}

__device__ __forceinline__ float es_phi_fw(float t) {
    const float z = t * (2.0f / SPREAD_W);
    const float a = 1.0f - z * z;
    return a > 0.0f ? expf(SPREAD_BETA * (sqrtf(a) - 1.0f)) : 0.0f;
}

__device__ __forceinline__ void cell_frac(float k, int64_t n, int64_t& cell, float& frac) {
    const double s = (double)k * (double)n;
    const double fl = floor(s);
    frac = (float)(s - fl);
    int64_t c = (int64_t)fl % n;
    if (c < 0) c += n;
    cell = c;
}

struct FwdGeom {
    int64_t n[3];
};

// one wavefront per sample; lane = kernel cell (ox, oy), z looped.  The
// sample's position, weights and frame are read once; its coils run in
// groups of FW_CB whose grid loads are independent of each other (several
// loads in flight per wavefront); a group is reduced over the 64 cells by one
// reduce-scatter (10 lane exchanges for 4 coils instead of 48).  Each
// (sample, coil) sums its K channels in channel order and its cells in a
// fixed tree.
constexpr int FW_CB = 4;  // the reduce-scatter below splits lane bits 5..3 over 4 coils x 2 parts
static_assert(FW_CB == 4, "the interpolation reduce-scatter is written for 4 coils");
template <int D>
__global__ void __launch_bounds__(256)
interp_kernel(const float* __restrict__ kc, const int* __restrict__ frame,
              const int64_t* __restrict__ sel, int64_t M, int64_t nsel, const float2* __restrict__ G,
              FwdGeom g, int K, int C, const void* __restrict__ U, int u_real, int64_t nt,
              float2* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int ox = (lane & 7) - 3, oy = (lane >> 3) - 3;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t Gc = g.n[0] * g.n[1] * g.n[2];
    for (int64_t j = wave0; j < M; j += nwaves) {
        int64_t cx, cy, cz = 0;
        float fx, fy, fz = 0.f;
        cell_frac(kc[j * D + 0], g.n[0], cx, fx);
        cell_frac(kc[j * D + 1], g.n[1], cy, fy);
        if (D == 3) cell_frac(kc[j * D + 2], g.n[2], cz, fz);
        const float wxy = es_phi_fw((float)ox - fx) * es_phi_fw((float)oy - fy);
        int64_t gx = (cx + ox) % g.n[0];
        if (gx < 0) gx += g.n[0];
        int64_t gy = (cy + oy) % g.n[1];
        if (gy < 0) gy += g.n[1];
        const int64_t off = gy * g.n[0] + gx;
        const int t = frame[j];
        const int64_t idx = sel ? sel[j] : j;
        for (int c0 = 0; c0 < C; c0 += FW_CB) {
            float2 acc[FW_CB];
#pragma unroll
            for (int cc = 0; cc < FW_CB; ++cc) acc[cc] = make_float2(0.f, 0.f);
            for (int k = 0; k < K; ++k) {
                float2 u;
                if (!U) u = make_float2(t == k ? 1.f : 0.f, 0.f);
                else if (u_real) u = make_float2(reinterpret_cast<const float*>(U)[t + (int64_t)k * nt], 0.f);
                else u = reinterpret_cast<const float2*>(U)[t + (int64_t)k * nt];
#pragma unroll
                for (int cc = 0; cc < FW_CB; ++cc) {
                    const int c = c0 + cc < C ? c0 + cc : C - 1;  // a spare slot repeats the last coil
                    const float2* Gk = G + ((int64_t)c * K + k) * Gc;
                    float2 s = make_float2(0.f, 0.f);
                    if (D == 2) {
                        const float2 v = Gk[off];
                        s = make_float2(wxy * v.x, wxy * v.y);
                    } else {
#pragma unroll
                        for (int oz = -3; oz <= 4; ++oz) {
                            int64_t gz = (cz + oz) % g.n[2];
                            if (gz < 0) gz += g.n[2];
                            const float wz = wxy * es_phi_fw((float)oz - fz);
                            const float2 v = Gk[gz * g.n[1] * g.n[0] + off];
                            s.x += wz * v.x;
                            s.y += wz * v.y;
                        }
                    }
                    cfma(acc[cc], u, s);
                }
            }
            // reduce-scatter over the 64 lanes: lane bits 5, 4 pick the coil
            // of the group, bit 3 the component, then bits 2..0 sum it
            float a[2 * FW_CB];
#pragma unroll
            for (int cc = 0; cc < FW_CB; ++cc) {
                a[2 * cc] = acc[cc].x;
                a[2 * cc + 1] = acc[cc].y;
            }
#pragma unroll
            for (int h = 2 * FW_CB / 2, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
                const bool up = (lane & bit) != 0;
#pragma unroll
                for (int i = 0; i < h; ++i) {
                    const float send = up ? a[i] : a[h + i];
                    const float keep = up ? a[h + i] : a[i];
                    a[i] = keep + __shfl_xor(send, bit, 64);
                }
            }
#pragma unroll
            for (int o = 4; o > 0; o >>= 1) a[0] += __shfl_xor(a[0], o, 64);
            const int cc = lane >> 4, comp = (lane >> 3) & 1;
            if ((lane & 7) == 0 && c0 + cc < C)
                reinterpret_cast<float*>(out)[2 * (idx + (int64_t)(c0 + cc) * nsel) + comp] = a[0];
        }
    }
}

void forward_model(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N, const void* x,
                   int K, const void* U, int u_real, int ncoils, const void* const* cmaps, void* out,
                   hipStream_t st) {
    MSR_REQUIRE(!traj->integer, MSR_ERR_ARGUMENT, "the forward model takes a non-Cartesian trajectory");
    MSR_REQUIRE(traj->d == d, MSR_ERR_ARGUMENT, "trajectory has %d dimensions, image %d", traj->d, d);
    MSR_REQUIRE(K >= 1, MSR_ERR_ARGUMENT, "K = %d", K);
    MSR_REQUIRE(U != nullptr || K == traj->nt, MSR_ERR_ARGUMENT,
                "U = I(Nt) needs K = Nt (%lld), got %d", (long long)traj->nt, K);
    const int C = ncoils >= 1 ? ncoils : 1;
    int64_t nos[3] = {1, 1, 1}, NN[3] = {1, 1, 1}, V = 1, Gos = 1;
    for (int a = 0; a < d; ++a) {
        NN[a] = N[a];
        nos[a] = 2 * N[a];
        MSR_REQUIRE(N[a] >= 1 && fft_len_supported(nos[a]), MSR_ERR_UNSUPPORTED,
                    "forward-model grid 2N[%d] = %lld has no FFT kernel in this build", a,
                    (long long)nos[a]);
        V *= N[a];
        Gos *= nos[a];
    }
    const int nch = C * K;
    // workspace: Z (nch*V) and two grid buffers
    DevBuf work;
    const size_t zb = sizeof(float2) * (size_t)nch * V, gb = sizeof(float2) * (size_t)nch * Gos;
    work.alloc(zb + 2 * gb);
    float2* Z = work.as<float2>();
    float2* A = reinterpret_cast<float2*>(reinterpret_cast<char*>(work.p) + zb);
    float2* B = reinterpret_cast<float2*>(reinterpret_cast<char*>(work.p) + zb + gb);
    // pack the coil maps (C device pointers) into one buffer when given
    DevBuf maps;
    const float2* mp = nullptr;
    if (ncoils >= 1 && cmaps) {
        maps.alloc(sizeof(float2) * (size_t)C * V);
        for (int c = 0; c < C; ++c) {
            MSR_REQUIRE(cmaps[c] != nullptr, MSR_ERR_ARGUMENT, "cmaps[%d] is NULL", c);
            MSR_HIP(hipMemcpyAsync(maps.as<float2>() + (size_t)c * V, cmaps[c], sizeof(float2) * V,
                                   hipMemcpyDeviceToDevice, st));
        }
        mp = maps.as<float2>();
    }
    int grid = (int)std::min<int64_t>(((int64_t)nch * V + 255) / 256, 16384);
    hipLaunchKernelGGL(expand_kernel, dim3(grid), dim3(256), 0, st, (const float2*)x, mp, C, K, V, Z);
    MSR_CHECK_LAUNCH();
    // x: centred placement + deapodisation, forward FFT over 2Nx  -> A [ch][z][y][qx]
    LinePass px;
    px.in = Z;
    px.out = A;
    px.nf = (int)nos[0];
    px.dir = -1;
    px.lin = (int)NN[0];
    px.isel = SEL_CENTER;
    px.win = deapod_table(ctx, (int)nos[0], (int)NN[0], SEL_CENTER);
    px.lout = (int)nos[0];
    px.sel = SEL_FIRST;
    px.nlines = (int64_t)nch * NN[1] * NN[2];
    px.in_pitch = NN[0];
    px.out_pitch = nos[0];
    run_rows(px, ctx, st);
    // y: A [ch][z][y][qx] -> B [ch][z][qy][qx]
    LinePass py;
    py.in = A;
    py.out = B;
    py.nf = (int)nos[1];
    py.dir = -1;
    py.lin = (int)NN[1];
    py.isel = SEL_CENTER;
    py.win = deapod_table(ctx, (int)nos[1], (int)NN[1], SEL_CENTER);
    py.lout = (int)nos[1];
    py.sel = SEL_FIRST;
    py.nb = (int64_t)nch * NN[2];
    py.ncols = nos[0];
    py.in_bstride = NN[1] * nos[0];
    py.in_stride = nos[0];
    py.out_bstride = nos[1] * nos[0];
    py.out_stride = nos[0];
    run_cols(py, ctx, st);
    const float2* Gfin = B;
    if (d == 3) {
        // z: B [ch][z][qy][qx] -> A [ch][qz][qy][qx]
        LinePass pz;
        pz.in = B;
        pz.out = A;
        pz.nf = (int)nos[2];
        pz.dir = -1;
        pz.lin = (int)NN[2];
        pz.isel = SEL_CENTER;
        pz.win = deapod_table(ctx, (int)nos[2], (int)NN[2], SEL_CENTER);
        pz.lout = (int)nos[2];
        pz.sel = SEL_FIRST;
        pz.nb = nch;
        pz.ncols = nos[1] * nos[0];
        pz.in_bstride = NN[2] * nos[1] * nos[0];
        pz.in_stride = nos[1] * nos[0];
        pz.out_bstride = Gos;
        pz.out_stride = nos[1] * nos[0];
        run_cols(pz, ctx, st);
        Gfin = A;
    }
    const int64_t nsel = traj->nsamp * traj->nt;
    if (traj->sel.p) fill_floats(out, 2 * nsel * C, 0.f, st);  // masked-out samples are 0
    FwdGeom g{{nos[0], nos[1], nos[2]}};
    const int64_t waves = traj->M;
    const int blocks = (int)std::min<int64_t>((waves * 64 + 255) / 256, 65536);
    if (d == 2)
        hipLaunchKernelGGL(interp_kernel<2>, dim3(blocks), dim3(256), 0, st, traj->kc.as<float>(),
                           traj->frame.as<int>(), traj->sel.p ? traj->sel.as<int64_t>() : nullptr,
                           traj->M, nsel, Gfin, g, K, C, U, u_real, traj->nt, (float2*)out);
    else
        hipLaunchKernelGGL(interp_kernel<3>, dim3(blocks), dim3(256), 0, st, traj->kc.as<float>(),
                           traj->frame.as<int>(), traj->sel.p ? traj->sel.as<int64_t>() : nullptr,
                           traj->M, nsel, Gfin, g, K, C, U, u_real, traj->nt, (float2*)out);
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipStreamSynchronize(st));  // workspace is freed on return
}

}  // namespace msr
