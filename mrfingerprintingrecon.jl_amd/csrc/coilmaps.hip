// Coil-map calibration front end (calculate_coil_maps and reconstruct_coilwise,
// ext/MRISubspaceReconCUDAExt/CoilMaps.jl): the calibration-region selection,
// the coil-wise reconstruction at K = 1 on the Toeplitz normal operator, and
// the calibration k-space that ESPIRiT (a third-party CPU package) takes.
//
// The coil-wise reconstruction solves one CG per coil on one operator.  Here
// the coils are operators sharing one kernel, solved together by the batch
// engine: every CG iteration is one launch per apply pass and per vector step
// for all coils, where the reference runs the coils' solves one after another.
#include <memory>
#include <vector>

#include "batch.hpp"
#include "fft.hpp"
#include "fftlines.hpp"
#include "normalop.hpp"

namespace msr {

void op_build_from_traj(msr_normalop* op, const msr_traj* traj, const void* U, int u_real,
                        hipStream_t st);  // kernel_build.hip
void backprojection(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N, const void* data,
                    int ncoils, int K, const void* U, int u_real, const void* const* cmaps,
                    int dcf_mode, const float* dcf, void* out, hipStream_t st, bool accumulate);
void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

// trj_out = trj .* scale; mask_out = all(|trj| .* scale .< 0.5) .& mask
// (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:17-20), Float32 products as there.
__global__ __launch_bounds__(256) void calib_select_kernel(const float* __restrict__ trj,
                                                           const uint8_t* __restrict__ mask,
                                                           int d, int64_t n, float s0, float s1,
                                                           float s2, float* __restrict__ trj_out,
                                                           uint8_t* __restrict__ mask_out) {
    const float s[3] = {s0, s1, s2};
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
        bool in = mask ? mask[j] != 0 : true;
        for (int a = 0; a < d; ++a) {
            const float v = trj[j * d + a];
            in = in && (fabsf(v) * s[a] < 0.5f);
            trj_out[j * d + a] = v * s[a];
        }
        mask_out[j] = in ? 1 : 0;
    }
}

void calib_select(int d, int64_t nsamp, int64_t nt, const float* trj, const uint8_t* mask,
                  const int64_t* N, const int64_t* Nc, float* trj_out, uint8_t* mask_out,
                  hipStream_t st) {
    MSR_REQUIRE(d == 2 || d == 3, MSR_ERR_ARGUMENT, "d = %d", d);
    MSR_REQUIRE(nsamp >= 1 && nt >= 1, MSR_ERR_ARGUMENT, "nsamp = %lld, nt = %lld",
                (long long)nsamp, (long long)nt);
    float s[3] = {1.f, 1.f, 1.f};
    for (int a = 0; a < d; ++a) {
        MSR_REQUIRE(N[a] >= 1 && Nc[a] >= 1, MSR_ERR_ARGUMENT, "img_shape[%d] = %lld, calib_size = %lld",
                    a, (long long)N[a], (long long)Nc[a]);
        int64_t m = Nc[a];
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        MSR_REQUIRE(m == 1, MSR_ERR_ASSERTION,
                    "calib_size has to be composed of the prime factors 2, 3, and 5 "
                    "(calib_size[%d] = %lld)", a, (long long)Nc[a]);
        s[a] = (float)((double)N[a] / (double)Nc[a]);  // cu() of img_shape ./ calib_size
    }
    const int64_t n = nsamp * nt;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(calib_select_kernel, dim3(grid), dim3(256), 0, st, trj, mask, d, n, s[0],
                       s[1], s[2], trj_out, mask_out);
    MSR_CHECK_LAUNCH();
}

// out[c][(z + nz/2) % nz][(y + ny/2) % ny][(x + nx/2) % nx] = in[c][z][y][x]
__global__ __launch_bounds__(256) void fftshift_kernel(const float2* __restrict__ in,
                                                       float2* __restrict__ out, int64_t total,
                                                       int nx, int ny, int nz) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const int64_t r = i / nx;
        const int y = (int)(r % ny);
        const int64_t r2 = r / ny;
        const int z = (int)(r2 % nz);
        const int64_t c = r2 / nz;
        const int64_t o = ((c * nz + (z + nz / 2) % nz) * ny + (y + ny / 2) % ny) * nx + (x + nx / 2) % nx;
        out[o] = in[i];
    }
}

// kbp = fftshift(fft(fftshift(x, dims)), dims) over the image axes of every
// coil image (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:22-25); x, out
// (C, [Nz,] Ny, Nx) ComplexF32, unscaled forward FFT.
void calib_kspace(msr_ctx* ctx, int d, const int64_t* Nc, int ncoils, const float2* x,
                  float2* out, hipStream_t st) {
    MSR_REQUIRE(d == 2 || d == 3, MSR_ERR_ARGUMENT, "d = %d", d);
    MSR_REQUIRE(ncoils >= 1, MSR_ERR_ARGUMENT, "ncoils = %d", ncoils);
    MSR_REQUIRE(x != out, MSR_ERR_ARGUMENT, "x and out must not alias");
    int64_t n[3] = {1, 1, 1};
    for (int a = 0; a < d; ++a) {
        n[a] = Nc[a];
        MSR_REQUIRE(fft_len_supported(n[a]), MSR_ERR_UNSUPPORTED,
                    "calibration k-space needs an FFT of length %lld", (long long)n[a]);
    }
    const int64_t V = n[0] * n[1] * n[2], total = V * ncoils;
    DevBuf tmp;
    tmp.alloc(sizeof(float2) * (size_t)total);
    float2* T = tmp.as<float2>();
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(fftshift_kernel, dim3(grid), dim3(256), 0, st, x, out, total, (int)n[0],
                       (int)n[1], (int)n[2]);
    MSR_CHECK_LAUNCH();
    auto line = [&](int a) {
        LinePass p;
        p.nf = (int)n[a];
        p.dir = -1;
        p.lin = p.lout = (int)n[a];
        p.isel = SEL_FIRST;
        p.sel = SEL_CENTER;  // output bin j holds frequency j - n/2 (fftshift)
        return p;
    };
    // x lines: out -> T
    LinePass px = line(0);
    px.in = out;
    px.out = T;
    px.nlines = (int64_t)ncoils * n[1] * n[2];
    px.in_pitch = px.out_pitch = n[0];
    run_rows(px, ctx, st);
    // y lines: T -> out
    LinePass py = line(1);
    py.in = T;
    py.out = out;
    py.nb = (int64_t)ncoils * n[2];
    py.ncols = n[0];
    py.in_bstride = py.out_bstride = n[1] * n[0];
    py.in_stride = py.out_stride = n[0];
    run_cols(py, ctx, st);
    if (d == 3) {  // z lines: out -> T -> out
        LinePass pz = line(2);
        pz.in = out;
        pz.out = T;
        pz.nb = ncoils;
        pz.ncols = n[1] * n[0];
        pz.in_bstride = pz.out_bstride = V;
        pz.in_stride = pz.out_stride = n[1] * n[0];
        run_cols(pz, ctx, st);
        MSR_HIP(hipMemcpyAsync(out, T, sizeof(float2) * (size_t)total, hipMemcpyDeviceToDevice, st));
    }
    MSR_HIP(hipStreamSynchronize(st));  // tmp is released on return
}

// reconstruct_coilwise (ext/MRISubspaceReconCUDAExt/CoilMaps.jl:56-77): the
// K = 1 operator of U[:, 1] (ones(Nt) when U is NULL), the per-coil
// backprojection with U[:, 1] (density_compensation = :none), then
// cg!(x_c, AᴴA, b_c; maxiter = niter, reltol = 0) for every coil c.
// out: (C, N...) ComplexF32 (Julia (N..., C)).
void reconstruct_coilwise(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N,
                          const void* data, int ncoils, const void* U, int u_real, int niter,
                          int prec, float2* out, hipStream_t st) {
    MSR_REQUIRE(traj->d == d, MSR_ERR_ARGUMENT, "trajectory has %d dimensions, img_shape %d",
                traj->d, d);
    MSR_REQUIRE(ncoils >= 1, MSR_ERR_ARGUMENT, "ncoils = %d", ncoils);
    MSR_REQUIRE(niter >= 0, MSR_ERR_ARGUMENT, "Niter_cg = %d", niter);
    MSR_REQUIRE(prec_valid(prec), MSR_ERR_ARGUMENT, "precision = %d", prec);
    DevBuf ones;
    if (U == nullptr) {
        ones.alloc(sizeof(float) * (size_t)traj->nt);
        fill_floats(ones.p, traj->nt, 1.f, st);
        U = ones.p;
        u_real = 1;
    }
    std::unique_ptr<msr_normalop> op(op_new(ctx, d, N, 1, u_real ? 1 : 0));
    op_set_cmaps(op.get(), 0, nullptr, st);
    op_build_from_traj(op.get(), traj, U, u_real, st);
    op_alloc_scratch(op.get());
    const int64_t V = op->V;
    DevBuf xbp;
    xbp.alloc(sizeof(float2) * (size_t)ncoils * V);
    backprojection(ctx, traj, d, N, data, ncoils, 1, U, u_real, nullptr, MSR_DCF_NONE, nullptr,
                   xbp.p, st, false);
    // one operator per coil on the shared kernel, solved in lockstep
    std::vector<std::unique_ptr<msr_normalop>> more;
    std::vector<msr_normalop*> ops{op.get()};
    for (int c = 1; c < ncoils; ++c) {
        more.emplace_back(op_on_kernel(ctx, d, op->N, 1, op->lam_real, op->nkm, op->lam,
                                       op->kmask, 0, nullptr, st));
        ops.push_back(more.back().get());
    }
    std::unique_ptr<msr_batch> bt(batch_create(ctx, ncoils, ops.data()));
    std::vector<float2*> xs(ncoils);
    std::vector<const float2*> bs(ncoils);
    for (int c = 0; c < ncoils; ++c) {
        xs[c] = out + (size_t)c * V;
        bs[c] = xbp.as<float2>() + (size_t)c * V;
    }
    batch_cg(bt.get(), xs.data(), bs.data(), niter, 0.0, 0.0, true, MSR_CG_DEVICE, prec, st);
    MSR_HIP(hipStreamSynchronize(st));  // the workspaces are released on return
}

}  // namespace msr
