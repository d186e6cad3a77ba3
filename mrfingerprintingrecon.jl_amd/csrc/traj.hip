// Library context (device, twiddle tables) and trajectory sample selection.
//
// Sample selection restates trj[:, sample_mask] with per-frame counts and
// offsets (src/NFFTNormalOp.jl:126-131, ext/MRISubspaceReconCUDAExt/
// BackProjection.jl:37-48,56): samples are kept frame-major, in order within
// each frame, so frame t's samples occupy [off_t, off_t + nsamp_t).
#include <cmath>
#include <vector>

#include "fft.hpp"
#include "internal.hpp"

msr_ctx::~msr_ctx() {
    for (auto& kv : twiddles) (void)hipFree(kv.second);
    for (auto& kv : twiddles64) (void)hipFree(kv.second);
    for (auto& kv : deapod) (void)hipFree(kv.second);
    for (auto& kv : deapod64) (void)hipFree(kv.second);
}

namespace msr {

void ensure_device(const msr_ctx* ctx) { MSR_HIP(hipSetDevice(ctx->device)); }

// exp(-2 pi i m / n), m < n, in double; exact at multiples of pi/2
static std::vector<double2> twiddles_f64(int n) {
    std::vector<double2> h(n);
    for (int m = 0; m < n; ++m) {
        // exp(-2 pi i m / n) in double, rounded once; exact at multiples of pi/2
        const int64_t q4 = 4LL * m;
        double c, s;
        if (q4 % n == 0) {
            const int quad = (int)(q4 / n) & 3;
            c = quad == 0 ? 1.0 : (quad == 2 ? -1.0 : 0.0);
            s = quad == 1 ? -1.0 : (quad == 3 ? 1.0 : 0.0);
        } else {
            const double a = -2.0 * M_PI * (double)m / (double)n;
            c = std::cos(a);
            s = std::sin(a);
        }
        h[m] = make_double2(c, s);
    }
    return h;
}

// Stage-ordered twiddles of the length-n plan (the layout fft_stage reads):
// for stage s >= 1, with NS the product of the earlier radices and R its
// radix, entry (NS - R0) + (r - 1) NS + k = exp(-2 pi i k r / (NS R)) for
// k < NS, 1 <= r < R.  Consecutive k are adjacent, so the threads of a stage
// read it without LDS bank conflicts.  n - R0 entries, padded to n.
static std::vector<double2> stage_twiddles_f64(int n) {
    const std::vector<double2> w = twiddles_f64(n);
    std::vector<double2> h(n, make_double2(0.0, 0.0));
    const bool ok = dispatch_line_len(n, [&]<int N>() {
        using Pl = FFTPlan<N>;
        int ns = Pl::R[0];
        for (int s = 1; s < Pl::NST; ++s) {
            const int R = Pl::R[s];
            for (int r = 1; r < R; ++r)
                for (int k = 0; k < ns; ++k)
                    h[(ns - Pl::R[0]) + (r - 1) * ns + k] = w[(int64_t)k * r * (N / (ns * R))];
            ns *= R;
        }
    });
    MSR_REQUIRE(ok, MSR_ERR_UNSUPPORTED, "no FFT plan for length %d", n);
    return h;
}

const float2* twiddle_table(msr_ctx* ctx, int n) {
    std::lock_guard<std::mutex> guard(ctx->lock);
    auto it = ctx->twiddles.find(n);
    if (it != ctx->twiddles.end()) return it->second;
    const std::vector<double2> h64 = stage_twiddles_f64(n);
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) h[m] = make_float2((float)h64[m].x, (float)h64[m].y);  // rounded once
    float2* d = nullptr;
    MSR_HIP(hipMalloc(&d, sizeof(float2) * n));
    MSR_HIP(hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    ctx->twiddles[n] = d;
    return d;
}

const double2* twiddle_table64(msr_ctx* ctx, int n) {
    std::lock_guard<std::mutex> guard(ctx->lock);
    auto it = ctx->twiddles64.find(n);
    if (it != ctx->twiddles64.end()) return it->second;
    const std::vector<double2> h = stage_twiddles_f64(n);
    double2* d = nullptr;
    MSR_HIP(hipMalloc(&d, sizeof(double2) * n));
    MSR_HIP(hipMemcpy(d, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    ctx->twiddles64[n] = d;
    return d;
}

// ------------------------------------------------------------ selection
__global__ void count_frames_kernel(const uint8_t* __restrict__ mask, int64_t nsamp,
                                    int64_t* __restrict__ counts) {
    const int64_t t = blockIdx.x;
    __shared__ int64_t s[4];
    int64_t c = 0;
    for (int64_t i = threadIdx.x; i < nsamp; i += blockDim.x) c += mask[t * nsamp + i] ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[t] = s[0] + s[1] + s[2] + s[3];
}

// offsets[t] = sum_{t' < t} counts[t'] (one block, sequential chunks)
__global__ void scan_frames_kernel(const int64_t* __restrict__ counts, int64_t nt,
                                   int64_t* __restrict__ offsets) {
    __shared__ int64_t s[1024];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nt; base += 1024) {
        const int64_t i = base + threadIdx.x;
        s[threadIdx.x] = i < nt ? counts[i] : 0;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int64_t v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += v;
            __syncthreads();
        }
        if (i < nt) offsets[i] = carry + s[threadIdx.x] - (i < nt ? counts[i] : 0);
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
}

// one block per frame: ordered compaction of that frame's masked samples
__global__ void compact_kernel(const float* __restrict__ trj, const uint8_t* __restrict__ mask,
                               int d, int64_t nsamp, const int64_t* __restrict__ offsets,
                               float* __restrict__ kc, int* __restrict__ frame,
                               int64_t* __restrict__ sel) {
    const int64_t t = blockIdx.x;
    __shared__ int s[256];
    __shared__ int64_t base;
    if (threadIdx.x == 0) base = offsets[t];
    __syncthreads();
    for (int64_t c0 = 0; c0 < nsamp; c0 += 256) {
        const int64_t i = c0 + threadIdx.x;
        const int m = (i < nsamp && mask[t * nsamp + i]) ? 1 : 0;
        s[threadIdx.x] = m;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            int v = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += v;
            __syncthreads();
        }
        if (m) {
            const int64_t pos = base + s[threadIdx.x] - 1;
            const int64_t src = t * nsamp + i;
            for (int a = 0; a < d; ++a) kc[pos * d + a] = trj[src * d + a];
            frame[pos] = (int)t;
            sel[pos] = src;
        }
        __syncthreads();
        if (threadIdx.x == 255) base += s[255];
        __syncthreads();
    }
}

__global__ void frame_index_kernel(int* __restrict__ frame, int64_t M, int64_t nsamp) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < M;
         j += (int64_t)gridDim.x * blockDim.x)
        frame[j] = (int)(j / nsamp);
}

msr_traj* traj_create(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const float* trj,
                      const uint8_t* mask, hipStream_t st) {
    MSR_REQUIRE(d == 2 || d == 3, MSR_ERR_ARGUMENT, "trajectory dimension %d, expected 2 or 3", d);
    MSR_REQUIRE(nsamp >= 1 && nt >= 1, MSR_ERR_ARGUMENT, "empty trajectory (%lld x %lld)",
                (long long)nsamp, (long long)nt);
    MSR_REQUIRE(trj != nullptr, MSR_ERR_ARGUMENT, "trj is NULL");
    auto tj = std::make_unique<msr_traj>();
    tj->ctx = ctx;
    tj->d = d;
    tj->nsamp = nsamp;
    tj->nt = nt;
    tj->nsamp_t.assign(nt, nsamp);
    if (!mask) {
        tj->M = nsamp * nt;
        tj->kc.alloc(sizeof(float) * d * tj->M);
        MSR_HIP(hipMemcpyAsync(tj->kc.p, trj, sizeof(float) * d * tj->M, hipMemcpyDeviceToDevice, st));
        tj->frame.alloc(sizeof(int) * tj->M);
        int grid = (int)std::min<int64_t>((tj->M + 255) / 256, 8192);
        hipLaunchKernelGGL(frame_index_kernel, dim3(grid), dim3(256), 0, st, tj->frame.as<int>(),
                           tj->M, nsamp);
        MSR_CHECK_LAUNCH();
        return tj.release();
    }
    DevBuf counts, offsets;
    counts.alloc(sizeof(int64_t) * nt);
    offsets.alloc(sizeof(int64_t) * nt);
    hipLaunchKernelGGL(count_frames_kernel, dim3(nt), dim3(256), 0, st, mask, nsamp,
                       counts.as<int64_t>());
    MSR_CHECK_LAUNCH();
    hipLaunchKernelGGL(scan_frames_kernel, dim3(1), dim3(1024), 0, st, counts.as<int64_t>(), nt,
                       offsets.as<int64_t>());
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipMemcpyAsync(tj->nsamp_t.data(), counts.p, sizeof(int64_t) * nt,
                           hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    int64_t M = 0;
    for (auto c : tj->nsamp_t) M += c;
    MSR_REQUIRE(M > 0, MSR_ERR_ASSERTION, "Sample mask removes all samples, cannot compute kernel.");
    tj->M = M;
    tj->kc.alloc(sizeof(float) * d * M);
    tj->frame.alloc(sizeof(int) * M);
    tj->sel.alloc(sizeof(int64_t) * M);
    hipLaunchKernelGGL(compact_kernel, dim3(nt), dim3(256), 0, st, trj, mask, d, nsamp,
                       offsets.as<int64_t>(), tj->kc.as<float>(), tj->frame.as<int>(),
                       tj->sel.as<int64_t>());
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipStreamSynchronize(st));  // counts/offsets are freed on return
    return tj.release();
}

}  // namespace msr
