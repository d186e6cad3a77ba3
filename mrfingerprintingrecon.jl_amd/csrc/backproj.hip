// Subspace backprojection: calculate_backprojection of the reference
// (ext/MRISubspaceReconCUDAExt/BackProjection.jl:1-79, src/BackProjection.jl:
// 23-102,169-202):
//   b_k[n] = sum_c conj(S_c[n]) sum_j w_j y_{j,c} conj(U[t_j, k]) exp(+2 pi i k_j.(n - N/2))
// as one type-1 NUFFT with C*K value channels: all channels spread in one
// pass over the sorted points (the reference runs K*C separate NUFFTs, each
// re-spreading every point), inverse FFTs on the sigma = 2 grid with the
// centred crop and deapodisation fused into the line passes, then the
// conjugate-coil combination.
#include <algorithm>

#include "fft.hpp"
#include "fftlines.hpp"
#include "spread.hpp"

namespace msr {

constexpr int CMAP_CHUNK = 32;
struct CmapPtrs {
    const float2* p[CMAP_CHUNK];
};

// out[k][v] (+)= sum_{c in chunk} conj(S_c[v]) img[(c0 + c)*K + k][v]
__global__ void combine_kernel(const float2* __restrict__ img, CmapPtrs maps, int c0, int nc,
                               int K, int64_t V, int accumulate, float2* __restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)K * V;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(e / V);
        const int64_t v = e - (int64_t)k * V;
        float2 acc = accumulate ? out[e] : make_float2(0.f, 0.f);
        for (int c = 0; c < nc; ++c)
            cfmac(acc, maps.p[c][v], img[((int64_t)(c0 + c) * K + k) * V + v]);
        out[e] = acc;
    }
}

void cart_backprojection(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N,
                         const void* data, int ncoils, int K, const void* U, int u_real,
                         const void* const* cmaps, void* out, hipStream_t st);  // cartesian.hip

void backprojection(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N, const void* data,
                    int ncoils, int K, const void* U, int u_real, const void* const* cmaps,
                    int dcf_mode, const float* dcf, void* out, hipStream_t st, bool accumulate) {
    if (traj->integer) {  // the Integer-trajectory method (src/BackProjection.jl:104-156)
        MSR_REQUIRE(dcf_mode == MSR_DCF_NONE && !accumulate, MSR_ERR_ARGUMENT,
                    "an integer (Cartesian) trajectory takes no density compensation");
        MSR_REQUIRE(ncoils >= 1 && K >= 1, MSR_ERR_ARGUMENT, "ncoils = %d, K = %d", ncoils, K);
        MSR_REQUIRE(U != nullptr || K == traj->nt, MSR_ERR_ARGUMENT,
                    "U = I(Nt) needs K = Nt (%lld), got %d", (long long)traj->nt, K);
        cart_backprojection(ctx, traj, d, N, data, ncoils, K, U, u_real, cmaps, out, st);
        return;
    }
    MSR_REQUIRE(traj->d == d, MSR_ERR_ARGUMENT,
                "`cmaps` contains %d image dimensions, yet the 1st dimension of trj is of "
                "length %d", d, traj->d);
    MSR_REQUIRE(ncoils >= 1, MSR_ERR_ARGUMENT, "ncoils = %d", ncoils);
    MSR_REQUIRE(K >= 1, MSR_ERR_ARGUMENT, "K = %d", K);
    MSR_REQUIRE(U != nullptr || K == traj->nt, MSR_ERR_ARGUMENT,
                "U = I(Nt) needs K = Nt (%lld), got %d", (long long)traj->nt, K);
    MSR_REQUIRE(dcf_mode >= 0 && dcf_mode <= 3, MSR_ERR_ERROR,
                "`density_compensation` can only be `:radial_3D`, `:radial_2D`, `:none`, or of "
                "type  `AbstractArray`");
    MSR_REQUIRE(dcf_mode != 3 || dcf != nullptr, MSR_ERR_ARGUMENT, "density_compensation array is NULL");
    int64_t nos[3] = {1, 1, 1}, NN[3] = {1, 1, 1};
    int64_t V = 1, Gos = 1;
    for (int a = 0; a < d; ++a) {
        MSR_REQUIRE(N[a] >= 1, MSR_ERR_ARGUMENT, "img_shape[%d] = %lld", a, (long long)N[a]);
        NN[a] = N[a];
        nos[a] = 2 * N[a];
        MSR_REQUIRE(fft_len_supported(nos[a]), MSR_ERR_UNSUPPORTED,
                    "backprojection grid 2N[%d] = %lld has no FFT kernel in this build", a,
                    (long long)nos[a]);
        V *= N[a];
        Gos *= nos[a];
    }
    const int nch = ncoils * K;
    // the plan's point streams and the workspace belong to the handle: hold
    // it for the call, and order this stream after the workspace's last user
    std::lock_guard<std::recursive_mutex> lock(traj->mu);
    if (traj->work_event && traj->work_stream != st)
        MSR_HIP(hipStreamWaitEvent(st, traj->work_event, 0));
    const SpreadPlan& plan = spread_plan(traj, nos, st);

    ValueSrc src{};
    src.kind = 0;
    src.data = (const float2*)data;
    src.sel = traj->sel.p ? traj->sel.as<int64_t>() : nullptr;
    src.nsel = traj->nsamp * traj->nt;
    src.dcf = dcf;
    src.dcf_mode = dcf_mode;
    src.K = K;
    src.U = U;
    src.u_real = u_real;
    src.nt = traj->nt;

    // workspace: spread grid G (nch*Gos), then the line-pass buffer H
    // (nch*Gos/2) in the space the spreading tiles used before the gather
    // kept on the trajectory between calls (calls on one trajectory are
    // stream-ordered; a handle is not re-entrant)
    DevBuf& work = traj->bp_work;
    const size_t gbytes = sizeof(float2) * (size_t)nch * Gos;
    const bool gather3 = d == 3 && spread_bp3_supported(traj);
    const size_t sbytes = d == 2 ? spread_bp_workspace_bytes(traj, plan, nch, ncoils, st)
                          : gather3 ? spread_bp3_workspace_bytes(traj, plan, nch, st) : 0;
    const size_t hbytes = std::max(sizeof(float2) * (size_t)nch * (Gos / 2), sbytes);
    if (work.bytes < gbytes + hbytes) {
        MSR_HIP(hipStreamSynchronize(st));  // earlier calls may still use the old buffer
        work.alloc(gbytes + hbytes);
    }
    float2* G = work.as<float2>();
    float2* H = reinterpret_cast<float2*>(reinterpret_cast<char*>(work.p) + gbytes);
    if (d == 2) spread_bp(traj, plan, src, nch, G, H, st);
    else if (gather3) spread_bp3(traj, plan, src, nch, G, H, st);
    else spread(traj, plan, src, nch, G, H, false, st);

    const bool combine = cmaps != nullptr;
    float2* img_out = combine ? G : (float2*)out;  // per-coil output is (C, K, N...)

    // x: inverse FFT of every row, centred crop to Nx, deapodise
    LinePass px;
    px.in = G;
    px.out = H;
    px.nf = (int)nos[0];
    px.dir = +1;
    px.lin = (int)nos[0];
    px.lout = (int)NN[0];
    px.sel = SEL_CENTER;
    px.w = deapod_table(ctx, (int)nos[0], (int)NN[0], SEL_CENTER);
    px.nlines = (int64_t)nch * nos[1] * nos[2];
    px.in_pitch = nos[0];
    px.out_pitch = NN[0];
    run_rows(px, ctx, st);

    if (d == 2) {
        LinePass py;
        py.in = H;
        py.out = img_out;
        py.nf = (int)nos[1];
        py.dir = +1;
        py.lin = (int)nos[1];
        py.lout = (int)NN[1];
        py.sel = SEL_CENTER;
        py.w = deapod_table(ctx, (int)nos[1], (int)NN[1], SEL_CENTER);
        py.nb = nch;
        py.ncols = NN[0];
        py.in_bstride = nos[1] * NN[0];
        py.in_stride = NN[0];
        py.out_bstride = NN[1] * NN[0];
        py.out_stride = NN[0];
        run_cols(py, ctx, st);
    } else {
        // y: H [ch][zos][yos][Nx] -> G [ch][zos][Ny][Nx]
        LinePass py;
        py.in = H;
        py.out = G;
        py.nf = (int)nos[1];
        py.dir = +1;
        py.lin = (int)nos[1];
        py.lout = (int)NN[1];
        py.sel = SEL_CENTER;
        py.w = deapod_table(ctx, (int)nos[1], (int)NN[1], SEL_CENTER);
        py.nb = (int64_t)nch * nos[2];
        py.ncols = NN[0];
        py.in_bstride = nos[1] * NN[0];
        py.in_stride = NN[0];
        py.out_bstride = NN[1] * NN[0];
        py.out_stride = NN[0];
        run_cols(py, ctx, st);
        // z: G [ch][zos][Ny][Nx] -> img [ch][Nz][Ny][Nx]
        float2* zout = combine ? H : (float2*)out;
        LinePass pz;
        pz.in = G;
        pz.out = zout;
        pz.nf = (int)nos[2];
        pz.dir = +1;
        pz.lin = (int)nos[2];
        pz.lout = (int)NN[2];
        pz.sel = SEL_CENTER;
        pz.w = deapod_table(ctx, (int)nos[2], (int)NN[2], SEL_CENTER);
        pz.nb = nch;
        pz.ncols = NN[1] * NN[0];
        pz.in_bstride = nos[2] * NN[1] * NN[0];
        pz.in_stride = NN[1] * NN[0];
        pz.out_bstride = V;
        pz.out_stride = NN[1] * NN[0];
        run_cols(pz, ctx, st);
        img_out = zout;
    }

    if (combine) {
        const int grid = (int)std::min<int64_t>(((int64_t)K * V + 255) / 256, 8192);
        for (int c0 = 0; c0 < ncoils; c0 += CMAP_CHUNK) {
            CmapPtrs mp{};
            const int nc = std::min(CMAP_CHUNK, ncoils - c0);
            for (int c = 0; c < nc; ++c) {
                MSR_REQUIRE(cmaps[c0 + c] != nullptr, MSR_ERR_ARGUMENT, "cmaps[%d] is NULL", c0 + c);
                mp.p[c] = (const float2*)cmaps[c0 + c];
            }
            hipLaunchKernelGGL(combine_kernel, dim3(grid), dim3(256), 0, st, img_out, mp, c0, nc, K,
                               V, (c0 > 0 || accumulate) ? 1 : 0, (float2*)out);
            MSR_CHECK_LAUNCH();
        }
    }
    // the next user of the workspace on another stream waits for this call
    if (!traj->work_event)
        MSR_HIP(hipEventCreateWithFlags(&traj->work_event, hipEventDisableTiming));
    MSR_HIP(hipEventRecord(traj->work_event, st));
    traj->work_stream = st;
}

// The reference's call sequence (ext/MRISubspaceReconCUDAExt/BackProjection.jl:
// 56-66): one single-channel type-1 NUFFT per (coefficient, coil), each
// spreading every sample, its image accumulated into b_k with conj(S_c).  The
// result equals backprojection(); this form times the reference's sequence on
// this library's kernels (the baseline the fused form is compared with).
void backprojection_channelwise(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N,
                                const void* data, int ncoils, int K, const void* U, int u_real,
                                const void* const* cmaps, int dcf_mode, const float* dcf, void* out,
                                hipStream_t st) {
    MSR_REQUIRE(!traj->integer, MSR_ERR_ARGUMENT,
                "the channel-wise backprojection takes a non-Cartesian trajectory");
    MSR_REQUIRE(U != nullptr, MSR_ERR_ARGUMENT, "the channel-wise backprojection needs a basis U");
    MSR_REQUIRE(cmaps != nullptr, MSR_ERR_ARGUMENT, "the channel-wise backprojection needs coil maps");
    MSR_REQUIRE(ncoils >= 1 && K >= 1, MSR_ERR_ARGUMENT, "ncoils = %d, K = %d", ncoils, K);
    int64_t V = 1;
    for (int a = 0; a < d; ++a) V *= N[a];
    const int64_t nsel = traj->nsamp * traj->nt;
    for (int k = 0; k < K; ++k) {
        const void* Uk = u_real ? (const void*)((const float*)U + (int64_t)k * traj->nt)
                                : (const void*)((const float2*)U + (int64_t)k * traj->nt);
        for (int c = 0; c < ncoils; ++c) {
            const void* data_c = (const float2*)data + (int64_t)c * nsel;
            backprojection(ctx, traj, d, N, data_c, 1, 1, Uk, u_real, cmaps + c, dcf_mode, dcf,
                           (float2*)out + (int64_t)k * V, st, c > 0);
        }
    }
}

}  // namespace msr
