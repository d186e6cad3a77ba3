// Measured choice of the fused 2D apply's launch knobs: the counterpart of
// the reference's measured FFT planning (plan_fft! with FFTW.MEASURE,
// src/NFFTNormalOp.jl:71-72,141), over the choices this apply has that leave
// its result unchanged bit for bit: the number of coil groups of pass B (each
// coil's columns are transformed and multiplied alone, whichever group runs
// them) and the block orders of passes A and C (the same blocks and
// arithmetic in another order).  Every candidate's partial sums are compared
// with the first candidate's, word for word: a knob that changed a bit would
// be a defect, and the tuner fails on it rather than keeping it.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>

#include <vector>

#include "normalop.hpp"

namespace msr {

// deterministic input in [-1, 1) + i [-1, 1) (a 64-bit mix of the index)
template <class AT>
__global__ __launch_bounds__(256) void tune_fill_kernel(AT* __restrict__ b, int64_t n) {
    using RT = real_t<AT>;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = (uint64_t)i * 0x9e3779b97f4a7c15ull + 0x632be59bd9b4e019ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        const RT re = (RT)((double)(z >> 40) / (double)(1ull << 23) - 1.0);
        const RT im = (RT)((double)(z & 0xffffffull) / (double)(1ull << 23) - 1.0);
        b[i] = AT{re, im};
    }
}

// *ndiff += number of 32-bit words in which a and b differ
__global__ __launch_bounds__(256) void tune_compare_kernel(const uint32_t* __restrict__ a,
                                                           const uint32_t* __restrict__ b,
                                                           int64_t n, int* __restrict__ ndiff) {
    int local = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        local += a[i] != b[i];
    if (local) atomicAdd(ndiff, local);
}

static unsigned tune_grid(int64_t n) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
}

void op_tune(msr_normalop* op, int am, int nrep, double* times, int* chosen, hipStream_t st) {
    MSR_REQUIRE(am >= AM_F32 && am <= AM_F64C, MSR_ERR_ARGUMENT, "apply mode %d", am);
    MSR_REQUIRE(nrep >= 0 && nrep <= 1000, MSR_ERR_ARGUMENT, "nrep = %d", nrep);
    if (times)
        for (int i = 0; i < TUNE_CONFIGS; ++i) times[i] = std::numeric_limits<double>::quiet_NaN();
    auto index_of = [](const ApplyKnobs& k) {
        int g = 0;
        while ((1 << g) < k.Gb && g < 3) ++g;
        return 4 * g + 2 * k.order_a + k.order_c;
    };
    const bool fused2d = op->d == 2 && !op->cart && op->fused;
    if (!fused2d) {
        if (chosen) *chosen = -1;
        return;
    }
    if (nrep == 0) {
        if (chosen) *chosen = index_of(op->knobs[am]);
        return;
    }
    const bool f64 = am != AM_F32;
    op_ensure_scratch(op, f64);
    const size_t es = f64 ? sizeof(double2) : sizeof(float2);
    const int64_t n = (int64_t)op->K * op->V;
    DevBuf b, ref, ndiff;
    b.alloc(es * (size_t)n);
    ref.alloc(op->partial.bytes);
    ndiff.alloc(sizeof(int));
    if (f64)
        hipLaunchKernelGGL(tune_fill_kernel<double2>, dim3(tune_grid(n)), dim3(256), 0, st,
                           b.as<double2>(), n);
    else
        hipLaunchKernelGGL(tune_fill_kernel<float2>, dim3(tune_grid(n)), dim3(256), 0, st,
                           b.as<float2>(), n);
    MSR_CHECK_LAUNCH();
    hipEvent_t e0, e1;
    MSR_HIP(hipEventCreate(&e0));
    MSR_HIP(hipEventCreate(&e1));
    const ApplyKnobs saved = op->knobs[am];
    const int64_t words = (int64_t)(op->partial.bytes / sizeof(uint32_t));
    double best[TUNE_CONFIGS];
    for (double& v : best) v = std::numeric_limits<double>::infinity();
    bool have_ref = false;
    try {
        // two rounds over the candidates; each keeps its faster round
        for (int round = 0; round < 2; ++round)
            for (int g = 0; g < 4; ++g) {
                const int Gb = 1 << g;
                if (Gb > op->C) continue;
                for (int oa = 0; oa < 2; ++oa)
                    for (int oc = 0; oc < 2; ++oc) {
                        ApplyKnobs k;
                        k.Gb = Gb;
                        k.order_a = oa;
                        k.order_c = oc;
                        op->knobs[am] = k;
                        op_passes(op, b.p, am, st, nullptr);  // warm-up, and the bits to compare
                        if (!have_ref) {
                            MSR_HIP(hipMemcpyAsync(ref.p, op->partial.p, op->partial.bytes,
                                                   hipMemcpyDeviceToDevice, st));
                            have_ref = true;
                        } else if (round == 0) {
                            int zero = 0, nd = 0;
                            MSR_HIP(hipMemcpyAsync(ndiff.p, &zero, sizeof(int),
                                                   hipMemcpyHostToDevice, st));
                            hipLaunchKernelGGL(tune_compare_kernel, dim3(tune_grid(words)),
                                               dim3(256), 0, st, ref.as<uint32_t>(),
                                               op->partial.as<uint32_t>(), words,
                                               ndiff.as<int>());
                            MSR_CHECK_LAUNCH();
                            MSR_HIP(hipMemcpyAsync(&nd, ndiff.p, sizeof(int),
                                                   hipMemcpyDeviceToHost, st));
                            MSR_HIP(hipStreamSynchronize(st));
                            MSR_REQUIRE(nd == 0, MSR_ERR_ASSERTION,
                                        "apply knobs (Gb %d, order A %d, order C %d) changed %d "
                                        "words of the result", Gb, oa, oc, nd);
                        }
                        MSR_HIP(hipEventRecord(e0, st));
                        for (int r = 0; r < nrep; ++r) op_passes(op, b.p, am, st, nullptr);
                        MSR_HIP(hipEventRecord(e1, st));
                        MSR_HIP(hipEventSynchronize(e1));
                        float ms = 0.f;
                        MSR_HIP(hipEventElapsedTime(&ms, e0, e1));
                        const int idx = index_of(k);
                        best[idx] = std::min(best[idx], (double)ms / nrep);
                    }
            }
    } catch (...) {
        op->knobs[am] = saved;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int arg = index_of(saved);
    for (int i = 0; i < TUNE_CONFIGS; ++i) {
        if (times && std::isfinite(best[i])) times[i] = best[i];
        if (best[i] < best[arg]) arg = i;
    }
    ApplyKnobs k;
    k.Gb = 1 << (arg / 4);
    k.order_a = (arg / 2) % 2;
    k.order_c = arg % 2;
    op->knobs[am] = k;
    remember_knobs(op, am, k);
    // a captured CG graph holds the launch shapes of the old knobs: capture
    // again on its next use
    if (op->graph) {
        MSR_HIP(hipStreamSynchronize(op->graph_stream));  // a replay may still run
        (void)hipGraphExecDestroy(op->graph);
        op->graph = nullptr;
        op->graph_key.clear();
    }
    if (chosen) *chosen = arg;
}

// ------------------------------------------------------------ tuning file
// Text file of the context's tuned knobs, the counterpart of FFTW's saved
// wisdom: a header line "msr-tune 1 <arch> <compute units>" naming the device
// the times were measured on, then one line per shape and apply mode:
// nx ny K C kernel_real identity_maps apply_mode Gb order_a order_c.
static std::string device_tag(int device) {
    hipDeviceProp_t p;
    MSR_HIP(hipGetDeviceProperties(&p, device));
    std::string arch = p.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    return arch + " " + std::to_string(p.multiProcessorCount);
}

void tune_save(msr_ctx* ctx, const char* path) {
    const std::string tag = device_tag(ctx->device);
    std::FILE* f = std::fopen(path, "w");
    MSR_REQUIRE(f != nullptr, MSR_ERR_IO, "%s: cannot open for writing", path);
    bool ok = std::fprintf(f, "msr-tune 1 %s\n", tag.c_str()) > 0;
    {
        std::lock_guard<std::mutex> guard(ctx->lock);
        for (const auto& kv : ctx->tuned) {
            const auto& k = kv.first;
            ok = ok && std::fprintf(f, "%lld %lld %lld %lld %lld %lld %lld %d %d %d\n",
                                    (long long)k[0], (long long)k[1], (long long)k[2],
                                    (long long)k[3], (long long)k[4], (long long)k[5],
                                    (long long)k[6], kv.second.Gb, kv.second.order_a,
                                    kv.second.order_c) > 0;
        }
    }
    ok = (std::fclose(f) == 0) && ok;
    MSR_REQUIRE(ok, MSR_ERR_IO, "%s: write failed", path);
}

int tune_load(msr_ctx* ctx, const char* path) {
    std::FILE* f = std::fopen(path, "r");
    MSR_REQUIRE(f != nullptr, MSR_ERR_IO, "%s: cannot open", path);
    char magic[16] = {0}, arch[64] = {0};
    int version = 0, cus = 0;
    const int nh = std::fscanf(f, "%15s %d %63s %d", magic, &version, arch, &cus);
    if (nh != 4 || std::strcmp(magic, "msr-tune") != 0 || version != 1) {
        std::fclose(f);
        fail(MSR_ERR_IO, "%s: not a tuning file", path);
    }
    const std::string tag = std::string(arch) + " " + std::to_string(cus);
    if (tag != device_tag(ctx->device)) {
        std::fclose(f);
        fail(MSR_ERR_ARGUMENT, "%s: measured on %s, this device is %s", path, tag.c_str(),
             device_tag(ctx->device).c_str());
    }
    std::vector<std::pair<std::array<int64_t, 7>, ApplyKnobs>> rows;
    long long v[7];
    int gb, oa, oc;
    while (std::fscanf(f, "%lld %lld %lld %lld %lld %lld %lld %d %d %d", &v[0], &v[1], &v[2],
                       &v[3], &v[4], &v[5], &v[6], &gb, &oa, &oc) == 10) {
        const bool valid = v[0] >= 1 && v[1] >= 1 && v[2] >= 1 && v[3] >= 1 && v[6] >= 0 &&
                           v[6] <= 2 && gb >= 1 && gb <= 8 && (gb & (gb - 1)) == 0 &&
                           gb <= v[3] && (oa == 0 || oa == 1) && (oc == 0 || oc == 1);
        if (!valid) {
            std::fclose(f);
            fail(MSR_ERR_IO, "%s: invalid entry", path);
        }
        ApplyKnobs k;
        k.Gb = gb;
        k.order_a = oa;
        k.order_c = oc;
        rows.push_back({{v[0], v[1], v[2], v[3], v[4], v[5], v[6]}, k});
    }
    const bool at_end = std::feof(f) != 0;
    std::fclose(f);
    MSR_REQUIRE(at_end, MSR_ERR_IO, "%s: unreadable entry", path);
    std::lock_guard<std::mutex> guard(ctx->lock);
    for (const auto& r : rows) ctx->tuned[r.first] = r.second;
    return (int)rows.size();
}

}  // namespace msr
