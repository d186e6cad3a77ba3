// Toeplitz kernel construction on the device: calculate_kmask_indcs and
// calculate_kernel_noncartesian of the reference
// (ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:75-191; src/NFFTNormalOp.jl:106-225).
//
// kmask: every cell of the 2N grid within the half-support-4 footprint of a
// sample, located as NonuniformFFTs.jl's sigma = 1 spreading locates it:
// base cell floor(mod(-k, 1) * 2N) (evaluated exactly in FP64), cells
// base-3 .. base+4 per dimension, periodic.  Sorted, 1-based.
//
// Kernel: for each packed pair (j <= k) the PSF
//   p_jk[r] = sum_m conj(U[t_m, j]) U[t_m, k] exp(+2 pi i k_m . r),  r on the 2N grid (FFT order)
// is one channel of a type-1 NUFFT (the P channels spread together onto the
// sigma = 2 grid 4N, inverse FFTs with FFT-order crop and deapodisation), then
// lambda_jk = FFT_2N(p_jk), stored densely with zeros off the kmask.  All of it
// runs in FP64 and is rounded once when stored.  Real U
// stores Re(lambda) (the reference's brfft path, src/NFFTNormalOp.jl:206-219,
// up to the PSF's Nyquist planes).
#include <algorithm>
#include <vector>

#include "fft.hpp"
#include "fftlines.hpp"
#include "spread.hpp"

namespace msr {

void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

struct GridN {
    int64_t n[3];
};

__global__ void kmask_mark_kernel(const float* __restrict__ kc, int d, int64_t M, GridN g,
                                  uint8_t* __restrict__ mask) {
    const int per = d == 2 ? 64 : 512;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < M * per;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / per;
        const int o = (int)(e - j * per);
        const int off[3] = {(o & 7) - 3, ((o >> 3) & 7) - 3, (o >> 6) - 3};
        int64_t lin = 0, mul = 1;
        for (int a = 0; a < d; ++a) {
            double t = -(double)kc[j * d + a];
            t = t - floor(t);
            int64_t cell = (int64_t)floor(t * (double)g.n[a]);
            if (cell > g.n[a] - 1) cell = g.n[a] - 1;
            int64_t c = (cell + off[a]) % g.n[a];
            if (c < 0) c += g.n[a];
            lin += c * mul;
            mul *= g.n[a];
        }
        mask[lin] = 1;
    }
}

__global__ void mask_count_kernel(const uint8_t* __restrict__ mask, int64_t G,
                                  int* __restrict__ counts) {
    __shared__ int s[16];
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    int v = (i < G && mask[i]) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0;
        for (int w = 0; w < 16; ++w) a += s[w];
        counts[blockIdx.x] = a;
    }
}

// exclusive scan (int64 output), single block, offs[nb] = total
__global__ void count_scan_kernel(const int* __restrict__ counts, int64_t nb,
                                  int64_t* __restrict__ offs) {
    __shared__ int64_t s[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int64_t v0 = i < nb ? counts[i] : 0;
        s[threadIdx.x] = v0;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int64_t v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
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

__global__ void mask_compact_kernel(const uint8_t* __restrict__ mask, int64_t G,
                                    const int64_t* __restrict__ offs, int64_t* __restrict__ out) {
    __shared__ int s[1024];
    const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    const int m = (i < G && mask[i]) ? 1 : 0;
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    if (m) out[offs[blockIdx.x] + s[threadIdx.x] - 1] = i + 1;  // 1-based
}

// Build the byte mask on grid n2 and the compacted 1-based kmask.
void build_kmask(const msr_traj* traj, const int64_t* n2, DevBuf& mask, DevBuf& kmask,
                 int64_t& nkm, hipStream_t st) {
    const int d = traj->d;
    int64_t G = 1;
    GridN g{{1, 1, 1}};
    for (int a = 0; a < d; ++a) {
        g.n[a] = n2[a];
        G *= n2[a];
    }
    mask.alloc(G);
    fill_floats(mask.p, G / 4, 0.f, st);
    if (G % 4) {
        // tail bytes (grid sizes here are even, G is a multiple of 4)
        fail(MSR_ERR_UNSUPPORTED, "grid size not a multiple of 4");
    }
    const int per = d == 2 ? 64 : 512;
    int grid = (int)std::min<int64_t>((traj->M * per + 255) / 256, 65536);
    hipLaunchKernelGGL(kmask_mark_kernel, dim3(grid), dim3(256), 0, st, traj->kc.as<float>(), d,
                       traj->M, g, mask.as<uint8_t>());
    MSR_CHECK_LAUNCH();
    const int64_t nb = ceil_div(G, 1024);
    DevBuf counts, offs;
    counts.alloc(sizeof(int) * nb);
    offs.alloc(sizeof(int64_t) * (nb + 1));
    hipLaunchKernelGGL(mask_count_kernel, dim3((unsigned)nb), dim3(1024), 0, st,
                       mask.as<uint8_t>(), G, counts.as<int>());
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(count_scan_kernel, dim3(1), dim3(1024), 0, st, counts.as<int>(), nb,
                       offs.as<int64_t>());
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipMemcpyAsync(&nkm, offs.as<int64_t>() + nb, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    kmask.alloc(sizeof(int64_t) * std::max<int64_t>(nkm, 1));
    hipLaunchKernelGGL(mask_compact_kernel, dim3((unsigned)nb), dim3(1024), 0, st,
                       mask.as<uint8_t>(), G, offs.as<int64_t>(), kmask.as<int64_t>());
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipStreamSynchronize(st));
}

void op_build_from_traj(msr_normalop* op, const msr_traj* traj, const void* U, int u_real,
                        hipStream_t st) {
    MSR_REQUIRE(!traj->integer, MSR_ERR_ARGUMENT,
                "NFFTNormalOp takes a non-Cartesian trajectory (FFTNormalOp takes integer ones)");
    msr_ctx* ctx = op->ctx;
    const int d = op->d;
    const int P = op->P;
    int64_t n2[3] = {1, 1, 1}, n4[3] = {1, 1, 1};
    int64_t G2 = 1, G4 = 1;
    for (int a = 0; a < d; ++a) {
        n2[a] = 2 * op->N[a];
        n4[a] = 4 * op->N[a];
        MSR_REQUIRE(fft_len_supported(n4[a]), MSR_ERR_UNSUPPORTED,
                    "kernel construction needs an FFT of length 4N[%d] = %lld", a,
                    (long long)n4[a]);
        G2 *= n2[a];
        G4 *= n4[a];
    }
    // kmask
    DevBuf mask;
    build_kmask(traj, n2, mask, *op->kmask, op->nkm, st);

    // pair tables (packed order)
    std::vector<int> pj(P), pk(P);
    for (int k = 0; k < op->K; ++k)
        for (int j = 0; j <= k; ++j) {
            const int p = j + k * (k + 1) / 2;
            pj[p] = j;
            pk[p] = k;
        }
    DevBuf pairs;
    pairs.alloc(sizeof(int) * 2 * P);
    MSR_HIP(hipMemcpyAsync(pairs.p, pj.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));
    MSR_HIP(hipMemcpyAsync(pairs.as<int>() + P, pk.data(), sizeof(int) * P, hipMemcpyHostToDevice, st));

    ValueSrc src{};
    src.kind = 1;
    src.U = U;
    src.u_real = u_real;
    src.nt = traj->nt;
    src.K = op->K;

    // The kernel is computed in FP64 (FP64 spreading accumulation, weights,
    // deapodisation and FFTs) and rounded once to ComplexF32 / Float32 when
    // stored: FP32 arithmetic here perturbs the operator by ~1e-6, which the
    // CG iterations of an indefinite truncated operator amplify beyond 1e-4.
    // Pairs are processed in chunks so the FP64 grids stay within a budget.
    const bool f64 = true;
    const size_t es = f64 ? sizeof(double2) : sizeof(float2);
    const SpreadPlan& plan = spread_plan(traj, n4, st);
    const size_t per_ch = es * (size_t)G4 + std::max(es * (size_t)(G4 / 2),
                                                     spread_workspace_bytes(plan, 1, f64));
    const size_t budget = (size_t)4 << 30;
    const int pc = (int)std::max<size_t>(1, std::min<size_t>((size_t)P, budget / per_ch));
    DevBuf A, B;
    A.alloc(es * (size_t)pc * G4);
    B.alloc(std::max(es * (size_t)pc * (G4 / 2), spread_workspace_bytes(plan, pc, f64)));

    const size_t ls = op->lam_real ? sizeof(float) : sizeof(float2);
    op->lam->alloc(ls * (size_t)P * G2);

    for (int p0 = 0; p0 < P; p0 += pc) {
        const int np = std::min(pc, P - p0);
        src.pair_j = pairs.as<int>() + p0;
        src.pair_k = pairs.as<int>() + P + p0;
        spread(traj, plan, src, np, A.p, B.p, f64, st);

        // x: inverse FFT over 4Nx, FFT-order crop to 2Nx, deapodise
        LinePass px;
        px.in = A.p;
        px.out = B.p;
        px.f64 = f64;
        px.nf = (int)n4[0];
        px.dir = +1;
        px.lin = (int)n4[0];
        px.lout = (int)n2[0];
        px.sel = SEL_FFTORDER;
        px.w64 = deapod_table64(ctx, (int)n4[0], (int)n2[0], SEL_FFTORDER);
        px.nlines = (int64_t)np * n4[1] * n4[2];
        px.in_pitch = n4[0];
        px.out_pitch = n2[0];
        run_rows(px, ctx, st);
        // B: [p][z4][y4][x2]
        LinePass py;
        py.in = B.p;
        py.out = A.p;
        py.f64 = f64;
        py.nf = (int)n4[1];
        py.dir = +1;
        py.lin = (int)n4[1];
        py.lout = (int)n2[1];
        py.sel = SEL_FFTORDER;
        py.w64 = deapod_table64(ctx, (int)n4[1], (int)n2[1], SEL_FFTORDER);
        py.nb = (int64_t)np * n4[2];
        py.ncols = n2[0];
        py.in_bstride = n4[1] * n2[0];
        py.in_stride = n2[0];
        py.out_bstride = n2[1] * n2[0];
        py.out_stride = n2[0];
        run_cols(py, ctx, st);
        // A: [p][z4][y2][x2]
        void* psf = A.p;
        void* other = B.p;
        if (d == 3) {
            LinePass pz;
            pz.in = A.p;
            pz.out = B.p;
            pz.f64 = f64;
            pz.nf = (int)n4[2];
            pz.dir = +1;
            pz.lin = (int)n4[2];
            pz.lout = (int)n2[2];
            pz.sel = SEL_FFTORDER;
            pz.w64 = deapod_table64(ctx, (int)n4[2], (int)n2[2], SEL_FFTORDER);
            pz.nb = np;
            pz.ncols = n2[1] * n2[0];
            pz.in_bstride = n4[2] * n2[1] * n2[0];
            pz.in_stride = n2[1] * n2[0];
            pz.out_bstride = G2;
            pz.out_stride = n2[1] * n2[0];
            run_cols(pz, ctx, st);
            psf = B.p;
            other = A.p;
        }
        // psf: [p][z2][y2][x2], FFT order.  Forward FFT over the 2N grid.
        LinePass fx;
        fx.in = psf;
        fx.out = other;
        fx.f64 = f64;
        fx.nf = (int)n2[0];
        fx.dir = -1;
        fx.lin = (int)n2[0];
        fx.lout = (int)n2[0];
        fx.sel = SEL_FIRST;
        fx.nlines = (int64_t)np * n2[1] * n2[2];
        fx.in_pitch = n2[0];
        fx.out_pitch = n2[0];
        run_rows(fx, ctx, st);
        LinePass fy;
        fy.in = other;
        fy.f64 = f64;
        fy.nf = (int)n2[1];
        fy.dir = -1;
        fy.lin = (int)n2[1];
        fy.lout = (int)n2[1];
        fy.sel = SEL_FIRST;
        fy.nb = (int64_t)np * n2[2];
        fy.ncols = n2[0];
        fy.in_bstride = n2[1] * n2[0];
        fy.in_stride = n2[0];
        if (d == 2) {
            // final pass stores the dense kernel: lam[(qx*P + p)*ny2 + qy]
            fy.lam = op->lam->p;
            fy.mask = mask.as<uint8_t>();
            fy.P = P;
            fy.p0 = p0;
            fy.lam_real = op->lam_real;
            fy.nx2 = n2[0];
            run_cols(fy, ctx, st);
        } else {
            fy.out = psf;
            fy.out_bstride = n2[1] * n2[0];
            fy.out_stride = n2[0];
            run_cols(fy, ctx, st);
            LinePass fz;
            fz.in = psf;
            fz.f64 = f64;
            fz.nf = (int)n2[2];
            fz.dir = -1;
            fz.lin = (int)n2[2];
            fz.lout = (int)n2[2];
            fz.sel = SEL_FIRST;
            fz.nb = np;
            fz.ncols = n2[1] * n2[0];
            fz.in_bstride = G2;
            fz.in_stride = n2[1] * n2[0];
            fz.lam = op->lam->p;
            fz.mask = mask.as<uint8_t>();
            fz.P = P;
            fz.p0 = p0;
            fz.lam_real = op->lam_real;
            fz.nx2 = n2[0];
            fz.ny2 = n2[1];
            fz.lam3d = 1;
            run_cols(fz, ctx, st);
        }
    }
    MSR_HIP(hipStreamSynchronize(st));  // temporaries are freed on return
}

}  // namespace msr
