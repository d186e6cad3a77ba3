// Toeplitz-kernel persistence and cache.
//
// The reference builds an operator either from the trajectory and basis
// (calculate_kernel_noncartesian, ext/MRISubspaceReconCUDAExt/NFFTNormalOp.jl:1-20)
// or from a precomputed packed kernel and kmask (:22-70).  A server that sees
// the same protocol (trajectory, sample mask, basis, image size) again should
// take the second route: this file gives the kernel an on-disk form, a key
// that names the protocol it was built from, and a cache of kernels by key
// (device memory, least recently used first out within a byte budget,
// optionally backed by a directory of kernel files).
//
// Kernel file (little-endian):
//   bytes 0..127   header (KernelFileHeader below)
//   then           kmask, nkm Int64 (1-based indices into the 2N grid)
//   then           packed kernel, (P, nkm) column-major: per kmask entry its
//                  P = K(K+1)/2 upper-triangle values, ComplexF32 (interleaved
//                  float pairs) or Float32 for a real basis
// The header and the payload each carry a 128-bit checksum (sum over 64-bit
// words w_i of splitmix64(w_i xor (i*c + s)) in two lanes with different c
// and s; splitmix64 is the finaliser of Steele, Lea and Flood's SplitMix
// generator with Vigna's constants).  The same word hash, computed on the
// device, keys the cache.
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <list>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "normalop.hpp"

namespace msr {

void op_build_from_traj(msr_normalop* op, const msr_traj* traj, const void* U, int u_real,
                        hipStream_t st);  // kernel_build.hip

// ------------------------------------------------------------------ hashing
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

constexpr uint64_t HASH_C0 = 0x9e3779b97f4a7c15ull, HASH_C1 = 0xc2b2ae3d27d4eb4full;
constexpr uint64_t HASH_S1 = 0x5851f42d4c957f2dull;

__host__ __device__ __forceinline__ void hash_word(uint64_t w, uint64_t i, uint64_t seed,
                                                   uint64_t& h0, uint64_t& h1) {
    h0 += mix64(w ^ (i * HASH_C0 + seed));
    h1 += mix64(w ^ (i * HASH_C1 + (seed ^ HASH_S1)));
}

static void hash_finish(uint64_t h0, uint64_t h1, uint64_t nbytes, uint64_t out[2]) {
    out[0] = mix64(h0 ^ mix64(nbytes));
    out[1] = mix64(h1 + mix64(nbytes ^ HASH_S1));
}

// Word i is bytes [8i, 8i+8) (the last one zero-padded); nbytes % 4 == 0.
static void hash_host_range(const uint8_t* p, uint64_t nbytes, uint64_t w0, uint64_t w1,
                            uint64_t seed, uint64_t& h0, uint64_t& h1) {
    for (uint64_t i = w0; i < w1; ++i) {
        uint64_t w = 0;
        const uint64_t off = 8 * i;
        std::memcpy(&w, p + off, std::min<uint64_t>(8, nbytes - off));
        hash_word(w, i, seed, h0, h1);
    }
}

static void hash_host(const void* data, uint64_t nbytes, uint64_t seed, uint64_t out[2]) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    const uint64_t nw = (nbytes + 7) / 8;
    uint64_t h0 = 0, h1 = 0;
    const int nth = nbytes > (64u << 20) ? 8 : 1;
    if (nth == 1) {
        hash_host_range(p, nbytes, 0, nw, seed, h0, h1);
    } else {  // the sum is order-free: split the words over threads
        std::vector<uint64_t> a(nth, 0), b(nth, 0);
        std::vector<std::thread> th;
        for (int t = 0; t < nth; ++t)
            th.emplace_back([&, t] {
                hash_host_range(p, nbytes, nw * t / nth, nw * (t + 1) / nth, seed, a[t], b[t]);
            });
        for (auto& x : th) x.join();
        for (int t = 0; t < nth; ++t) {
            h0 += a[t];
            h1 += b[t];
        }
    }
    hash_finish(h0, h1, nbytes, out);
}

constexpr int HASH_THREADS = 256, HASH_BLOCKS = 1024;

// Per-block lane sums of hash_word over the 64-bit words of a device buffer
// read as 32-bit words (n32 of them).
__global__ __launch_bounds__(HASH_THREADS) void hash_kernel(const uint32_t* __restrict__ p,
                                                            int64_t n32, uint64_t seed,
                                                            uint64_t* __restrict__ part) {
    __shared__ uint64_t s0[HASH_THREADS], s1[HASH_THREADS];
    const int t = threadIdx.x;
    const int64_t nw = (n32 + 1) / 2;
    uint64_t h0 = 0, h1 = 0;
    for (int64_t i = (int64_t)blockIdx.x * HASH_THREADS + t; i < nw;
         i += (int64_t)gridDim.x * HASH_THREADS) {
        const uint64_t lo = p[2 * i];
        const uint64_t hi = 2 * i + 1 < n32 ? p[2 * i + 1] : 0u;
        hash_word(lo | (hi << 32), (uint64_t)i, seed, h0, h1);
    }
    s0[t] = h0;
    s1[t] = h1;
    __syncthreads();
    for (int s = HASH_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) {
            s0[t] += s0[t + s];
            s1[t] += s1[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

// Hashes of nbuf device buffers (one partial-sum launch each, one read-back).
static void hash_device(const void* const* bufs, const uint64_t* nbytes, int nbuf, uint64_t seed,
                        hipStream_t st, uint64_t (*out)[2]) {
    DevBuf part;
    part.alloc(sizeof(uint64_t) * 2 * HASH_BLOCKS * nbuf);
    std::vector<int> blocks(nbuf);
    for (int b = 0; b < nbuf; ++b) {
        MSR_REQUIRE(nbytes[b] % 4 == 0, MSR_ERR_ARGUMENT, "hashed buffer of %llu bytes",
                    (unsigned long long)nbytes[b]);
        const int64_t n32 = (int64_t)(nbytes[b] / 4), nw = (n32 + 1) / 2;
        blocks[b] = (int)std::max<int64_t>(1, std::min<int64_t>(HASH_BLOCKS, (nw + HASH_THREADS - 1) / HASH_THREADS));
        hipLaunchKernelGGL(hash_kernel, dim3(blocks[b]), dim3(HASH_THREADS), 0, st,
                           (const uint32_t*)bufs[b], n32, seed,
                           part.as<uint64_t>() + (size_t)2 * HASH_BLOCKS * b);
        MSR_CHECK_LAUNCH();
    }
    std::vector<uint64_t> h((size_t)2 * HASH_BLOCKS * nbuf);
    MSR_HIP(hipMemcpyAsync(h.data(), part.p, part.bytes, hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < nbuf; ++b) {
        uint64_t h0 = 0, h1 = 0;
        for (int k = 0; k < blocks[b]; ++k) {
            h0 += h[(size_t)2 * HASH_BLOCKS * b + 2 * k];
            h1 += h[(size_t)2 * HASH_BLOCKS * b + 2 * k + 1];
        }
        hash_finish(h0, h1, nbytes[b], out[b]);
    }
}

// ---------------------------------------------------------------- protocol key
// The kernel depends on the image size, K, the basis (values and whether it
// is real), the masked trajectory points and their frames (the per-frame
// sample counts); the key hashes exactly these, and the library version.
void kernel_key(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N, int K,
                const void* U, int u_real, hipStream_t st, uint64_t key[2]) {
    (void)ctx;
    MSR_REQUIRE(traj->d == d, MSR_ERR_ARGUMENT, "trajectory has %d dimensions, img_shape %d",
                traj->d, d);
    MSR_REQUIRE(K >= 1, MSR_ERR_ARGUMENT, "K = %d", K);
    int64_t meta[10] = {msr_version(), d, N[0], N[1], d == 3 ? N[2] : 1, K, u_real ? 1 : 0,
                        traj->M, traj->nt, 0};
    const void* bufs[2] = {traj->kc.p, U};
    const uint64_t nb[2] = {(uint64_t)traj->M * d * sizeof(float),
                            (uint64_t)traj->nt * K * (u_real ? sizeof(float) : sizeof(float2))};
    uint64_t hd[2][2], hm[2], hf[2];
    hash_device(bufs, nb, 2, 0x6b6d61736b6b6579ull, st, hd);
    hash_host(meta, sizeof(meta), 1, hm);
    hash_host(traj->nsamp_t.data(), sizeof(int64_t) * traj->nsamp_t.size(), 2, hf);
    for (int l = 0; l < 2; ++l)
        key[l] = mix64(mix64(mix64(hm[l] ^ hd[0][l]) + hf[l]) ^ hd[1][l]);
}

// ---------------------------------------------------------------- kernel file
struct KernelFileHeader {
    char magic[8];            // "MSRTOEPL"
    uint32_t version;         // 1
    uint32_t header_bytes;    // 128
    uint32_t d, K, P, lambda_real;
    int64_t N[3];             // image size, 1 beyond d
    int64_t nkm;
    uint64_t key[2];          // protocol key, 0 if unknown
    uint64_t payload_bytes;   // 8 nkm + P nkm (8 or 4)
    uint64_t payload_hash[2];
    uint64_t header_hash;     // hash_host of the 104 bytes before it, seed 3
    uint8_t reserved[16];
};
static_assert(sizeof(KernelFileHeader) == 128, "kernel file header is 128 bytes");
static_assert(offsetof(KernelFileHeader, header_hash) == 104, "header layout");
static const char KMAGIC[8] = {'M', 'S', 'R', 'T', 'O', 'E', 'P', 'L'};

static uint64_t header_hash(const KernelFileHeader& h) {
    uint64_t o[2];
    hash_host(&h, offsetof(KernelFileHeader, header_hash), 3, o);
    return o[0];
}

struct File {
    FILE* f = nullptr;
    explicit File(FILE* x) : f(x) {}
    ~File() {
        if (f) fclose(f);
    }
};

void kernel_save(const msr_normalop* op, const char* path, const uint64_t* key, hipStream_t st) {
    MSR_REQUIRE(!op->cart, MSR_ERR_UNSUPPORTED,
                "kernel files hold the non-Cartesian (packed Toeplitz) kernel only");
    const size_t es = op->lam_real ? sizeof(float) : sizeof(float2);
    const size_t kb = sizeof(int64_t) * (size_t)op->nkm, lb = es * (size_t)op->P * op->nkm;
    std::vector<uint8_t> payload(kb + lb);
    if (op->nkm > 0) {
        DevBuf dl, dk;
        dl.alloc(lb);
        dk.alloc(kb);
        op_export(op, dl.p, dk.as<int64_t>(), st);
        MSR_HIP(hipMemcpyAsync(payload.data(), dk.p, kb, hipMemcpyDeviceToHost, st));
        MSR_HIP(hipMemcpyAsync(payload.data() + kb, dl.p, lb, hipMemcpyDeviceToHost, st));
        MSR_HIP(hipStreamSynchronize(st));
    }
    KernelFileHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, KMAGIC, 8);
    h.version = 1;
    h.header_bytes = sizeof(h);
    h.d = op->d;
    h.K = op->K;
    h.P = op->P;
    h.lambda_real = op->lam_real;
    for (int i = 0; i < 3; ++i) h.N[i] = op->N[i];
    h.nkm = op->nkm;
    if (key) {
        h.key[0] = key[0];
        h.key[1] = key[1];
    }
    h.payload_bytes = payload.size();
    hash_host(payload.data(), payload.size(), 4, h.payload_hash);
    h.header_hash = header_hash(h);
    // write beside the target and rename, so a reader never sees a partial file
    static std::atomic<uint64_t> serial{0};
    const std::string tmp = std::string(path) + ".tmp." + std::to_string((long long)getpid()) +
                            "." + std::to_string((unsigned long long)serial++);
    {
        File f(fopen(tmp.c_str(), "wb"));
        MSR_REQUIRE(f.f, MSR_ERR_IO, "cannot open %s for writing", tmp.c_str());
        bool ok = fwrite(&h, sizeof(h), 1, f.f) == 1;
        if (ok && !payload.empty()) ok = fwrite(payload.data(), payload.size(), 1, f.f) == 1;
        ok = ok && fflush(f.f) == 0;
        const int c = fclose(f.f);
        f.f = nullptr;
        if (!ok || c != 0) {
            unlink(tmp.c_str());
            fail(MSR_ERR_IO, "writing %s failed", tmp.c_str());
        }
    }
    if (rename(tmp.c_str(), path) != 0) {
        unlink(tmp.c_str());
        fail(MSR_ERR_IO, "cannot rename %s to %s", tmp.c_str(), path);
    }
}

// Reads and validates a kernel file's header (and, if payload != nullptr,
// its payload with the checksum).  Host only.
static KernelFileHeader read_kernel_file(const char* path, std::vector<uint8_t>* payload,
                                         bool verify) {
    File f(fopen(path, "rb"));
    MSR_REQUIRE(f.f, MSR_ERR_IO, "cannot open kernel file %s", path);
    MSR_REQUIRE(fseeko(f.f, 0, SEEK_END) == 0, MSR_ERR_IO, "cannot seek in %s", path);
    const int64_t size = (int64_t)ftello(f.f);
    MSR_REQUIRE(fseeko(f.f, 0, SEEK_SET) == 0, MSR_ERR_IO, "cannot seek in %s", path);
    KernelFileHeader h;
    MSR_REQUIRE(size >= (int64_t)sizeof(h) && fread(&h, sizeof(h), 1, f.f) == 1, MSR_ERR_IO,
                "%s is too short for a kernel file header (%lld bytes)", path, (long long)size);
    MSR_REQUIRE(std::memcmp(h.magic, KMAGIC, 8) == 0, MSR_ERR_IO, "%s is not a kernel file", path);
    MSR_REQUIRE(h.header_hash == header_hash(h), MSR_ERR_IO, "%s: header checksum mismatch",
                path);
    MSR_REQUIRE(h.version == 1 && h.header_bytes == sizeof(h), MSR_ERR_UNSUPPORTED,
                "%s: kernel file version %u is not supported", path, h.version);
    MSR_REQUIRE((h.d == 2 || h.d == 3) && h.K >= 1 && h.K <= (uint32_t)MAX_COEFFS && h.P == h.K * (h.K + 1) / 2 &&
                    h.lambda_real <= 1,
                MSR_ERR_IO, "%s: inconsistent header (d %u, K %u, P %u)", path, h.d, h.K, h.P);
    int64_t G = 1;
    for (int i = 0; i < 3; ++i) {
        MSR_REQUIRE(h.N[i] >= 1 && h.N[i] <= (1 << 16) && (i < (int)h.d || h.N[i] == 1),
                    MSR_ERR_IO, "%s: inconsistent header (N[%d] = %lld)", path, i,
                    (long long)h.N[i]);
        if (i < (int)h.d) G *= 2 * h.N[i];
    }
    MSR_REQUIRE(h.nkm >= 0 && h.nkm <= G, MSR_ERR_IO, "%s: nkm = %lld outside 0..%lld", path,
                (long long)h.nkm, (long long)G);
    const uint64_t want = (uint64_t)h.nkm * (8 + (uint64_t)h.P * (h.lambda_real ? 4 : 8));
    MSR_REQUIRE(h.payload_bytes == want, MSR_ERR_IO, "%s: payload size %llu, expected %llu", path,
                (unsigned long long)h.payload_bytes, (unsigned long long)want);
    MSR_REQUIRE((uint64_t)size == sizeof(h) + want, MSR_ERR_IO,
                "%s: file has %lld bytes, header says %llu", path, (long long)size,
                (unsigned long long)(sizeof(h) + want));
    if (payload || verify) {
        std::vector<uint8_t> local;
        std::vector<uint8_t>& buf = payload ? *payload : local;
        buf.resize(want);
        MSR_REQUIRE(want == 0 || fread(buf.data(), want, 1, f.f) == 1, MSR_ERR_IO,
                    "%s: short read", path);
        uint64_t ph[2];
        hash_host(buf.data(), want, 4, ph);
        MSR_REQUIRE(ph[0] == h.payload_hash[0] && ph[1] == h.payload_hash[1], MSR_ERR_IO,
                    "%s: payload checksum mismatch (corrupt kernel file)", path);
    }
    return h;
}

void kernel_file_info(const char* path, int verify, int* d, int64_t* N, int* K, int* lam_real,
                      int64_t* nkm, uint64_t* key) {
    const KernelFileHeader h = read_kernel_file(path, nullptr, verify != 0);
    if (d) *d = (int)h.d;
    if (N)
        for (int i = 0; i < 3; ++i) N[i] = h.N[i];
    if (K) *K = (int)h.K;
    if (lam_real) *lam_real = (int)h.lambda_real;
    if (nkm) *nkm = h.nkm;
    if (key) {
        key[0] = h.key[0];
        key[1] = h.key[1];
    }
}

msr_normalop* kernel_load(msr_ctx* ctx, const char* path, const uint64_t* expect_key,
                          int ncoils, const void* const* cmaps, hipStream_t st) {
    std::vector<uint8_t> payload;
    const KernelFileHeader h = read_kernel_file(path, &payload, true);
    if (expect_key)
        MSR_REQUIRE(h.key[0] == expect_key[0] && h.key[1] == expect_key[1], MSR_ERR_ARGUMENT,
                    "%s holds the kernel of another protocol (key %016llx%016llx)", path,
                    (unsigned long long)h.key[0], (unsigned long long)h.key[1]);
    std::unique_ptr<msr_normalop> op(op_new(ctx, (int)h.d, h.N, (int)h.K, (int)h.lambda_real));
    op_set_cmaps(op.get(), ncoils, cmaps, st);
    DevBuf dev;
    dev.alloc(std::max<size_t>(payload.size(), 8));
    if (!payload.empty())
        MSR_HIP(hipMemcpyAsync(dev.p, payload.data(), payload.size(), hipMemcpyHostToDevice, st));
    const int64_t* km = dev.as<int64_t>();
    op_load_packed(op.get(), km + h.nkm, km, h.nkm, st);  // bounds-checks kmask; synchronises
    op_alloc_scratch(op.get());
    MSR_HIP(hipStreamSynchronize(st));
    return op.release();
}

// A new operator on an existing kernel (shared device memory).
msr_normalop* op_on_kernel(msr_ctx* ctx, int d, const int64_t* N, int K, int lam_real,
                           int64_t nkm, const std::shared_ptr<DevBuf>& lam,
                           const std::shared_ptr<DevBuf>& kmask, int ncoils,
                           const void* const* cmaps, hipStream_t st, int cart) {
    std::unique_ptr<msr_normalop> op(op_new(ctx, d, N, K, lam_real, cart));
    op->nkm = nkm;
    op->lam = lam;
    op->kmask = kmask;
    op_set_cmaps(op.get(), ncoils, cmaps, st);
    op_alloc_scratch(op.get());
    MSR_HIP(hipStreamSynchronize(st));
    return op.release();
}

}  // namespace msr

// ------------------------------------------------------------------ cache
struct msr_kernel_cache {
    struct Entry {
        uint64_t key[2];
        int d, K, lam_real;
        int64_t N[3], nkm;
        std::shared_ptr<msr::DevBuf> lam, kmask;
        size_t bytes;
    };
    msr_ctx* ctx = nullptr;
    size_t budget = 0;
    std::string dir;             // empty: memory only
    std::list<Entry> lru;        // most recently used first
    size_t bytes = 0;
    int64_t counts[3] = {0, 0, 0};  // MSR_KERNEL_BUILT / _MEMORY / _DISK
    std::mutex lock;
};

namespace msr {

msr_kernel_cache* kernel_cache_create(msr_ctx* ctx, size_t budget, const char* dir) {
    auto* c = new msr_kernel_cache;
    c->ctx = ctx;
    c->budget = budget;
    if (dir && dir[0]) {
        struct stat sb;
        MSR_REQUIRE(stat(dir, &sb) == 0 && S_ISDIR(sb.st_mode), MSR_ERR_IO,
                    "kernel cache directory %s does not exist", dir);
        c->dir = dir;
    }
    return c;
}

static std::string key_hex(const uint64_t key[2]) {
    char b[40];
    snprintf(b, sizeof(b), "%016llx%016llx", (unsigned long long)key[0],
             (unsigned long long)key[1]);
    return b;
}

static void cache_insert(msr_kernel_cache* c, const uint64_t key[2], const msr_normalop* op) {
    msr_kernel_cache::Entry e;
    e.key[0] = key[0];
    e.key[1] = key[1];
    e.d = op->d;
    e.K = op->K;
    e.lam_real = op->lam_real;
    for (int i = 0; i < 3; ++i) e.N[i] = op->N[i];
    e.nkm = op->nkm;
    e.lam = op->lam;
    e.kmask = op->kmask;
    e.bytes = op_kernel_bytes(op);
    if (e.bytes > c->budget) return;  // too large to keep: the caller's operator still has it
    while (!c->lru.empty() && c->bytes + e.bytes > c->budget) {
        c->bytes -= c->lru.back().bytes;
        c->lru.pop_back();  // freed once no operator uses it
    }
    c->bytes += e.bytes;
    c->lru.push_front(std::move(e));
}

msr_normalop* kernel_cache_get(msr_kernel_cache* c, const msr_traj* traj, int d, const int64_t* N,
                               int K, const void* U, int u_real, int ncoils,
                               const void* const* cmaps, hipStream_t st, int* source,
                               uint64_t* key_out) {
    std::lock_guard<std::mutex> g(c->lock);
    uint64_t key[2];
    kernel_key(c->ctx, traj, d, N, K, U, u_real, st, key);
    if (key_out) {
        key_out[0] = key[0];
        key_out[1] = key[1];
    }
    for (auto it = c->lru.begin(); it != c->lru.end(); ++it) {
        if (it->key[0] != key[0] || it->key[1] != key[1]) continue;
        c->lru.splice(c->lru.begin(), c->lru, it);  // most recently used
        const auto& e = c->lru.front();
        msr_normalop* op = op_on_kernel(c->ctx, e.d, e.N, e.K, e.lam_real, e.nkm, e.lam, e.kmask,
                                        ncoils, cmaps, st);
        c->counts[MSR_KERNEL_MEMORY]++;
        if (source) *source = MSR_KERNEL_MEMORY;
        return op;
    }
    const std::string file = c->dir.empty() ? std::string() : c->dir + "/" + key_hex(key) + ".msrk";
    if (!file.empty() && access(file.c_str(), R_OK) == 0) {
        try {
            std::unique_ptr<msr_normalop> op(kernel_load(c->ctx, file.c_str(), key, ncoils, cmaps, st));
            cache_insert(c, key, op.get());
            c->counts[MSR_KERNEL_DISK]++;
            if (source) *source = MSR_KERNEL_DISK;
            return op.release();
        } catch (const Error& e) {
            // a corrupt or foreign file is rebuilt and overwritten below
            if (e.code != MSR_ERR_IO && e.code != MSR_ERR_ARGUMENT && e.code != MSR_ERR_UNSUPPORTED)
                throw;
        }
    }
    std::unique_ptr<msr_normalop> op(op_new(c->ctx, d, N, K, u_real ? 1 : 0));
    op_set_cmaps(op.get(), ncoils, cmaps, st);
    op_build_from_traj(op.get(), traj, U, u_real, st);
    op_alloc_scratch(op.get());
    MSR_HIP(hipStreamSynchronize(st));
    if (!file.empty()) kernel_save(op.get(), file.c_str(), key, st);
    cache_insert(c, key, op.get());
    c->counts[MSR_KERNEL_BUILT]++;
    if (source) *source = MSR_KERNEL_BUILT;
    return op.release();
}

void kernel_cache_stats(msr_kernel_cache* c, int64_t* entries, size_t* bytes, int64_t* counts) {
    std::lock_guard<std::mutex> g(c->lock);
    if (entries) *entries = (int64_t)c->lru.size();
    if (bytes) *bytes = c->bytes;
    if (counts)
        for (int i = 0; i < 3; ++i) counts[i] = c->counts[i];
}

void kernel_cache_destroy(msr_kernel_cache* c) { delete c; }

msr_ctx* cache_ctx(const msr_kernel_cache* c) { return c->ctx; }

}  // namespace msr
