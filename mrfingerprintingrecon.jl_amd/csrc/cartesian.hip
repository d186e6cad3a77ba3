// Cartesian subspace normal operator and backprojection: FFTNormalOp
// (src/FFTNormalOp.jl) and calculate_backprojection for an integer trajectory
// (src/BackProjection.jl:104-156).  The reference has these on the CPU only.
//
// Samples sit on the N grid itself: sample k (integer, 1..N centred) falls in
// cell mod(k - N/2 - 1, N) of the FFT-ordered grid (the ifftshift of
// src/FFTNormalOp.jl:89).  Kernel (full K x K per cell, cell-major,
// lam[q * K*K + j + K*k] = Julia Λ[j, k, q]):
//   Λ[j, k, q] = sum over kept samples of frame t in cell q of conj(U[t,j]) U[t,k]
// accumulated per cell in FP64 over the cell's samples in sample order (a
// stable cell sort), rounded once; kmask = cells with Λ[1,1] != 0.
// Apply on ping-pong grids (C, K, N...):
//   expand  T = S_c b_k
//   FFT along x, y (, z)                          (generic line passes)
//   multiply per cell, all coils, kernel read once
//   inverse FFT along z, y, x
//   combine partial_k = sum_c conj(S_c) T_ck      (1/V applied by the epilogue)
#include <hipcub/hipcub.hpp>

#include <memory>
#include <vector>

#include "fft.hpp"
#include "fftlines.hpp"
#include "normalop.hpp"

namespace msr {

msr_traj* traj_create(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const float* trj,
                      const uint8_t* mask, hipStream_t st);  // traj.hip
void fill_floats(void* p, int64_t n, float v, hipStream_t st);  // normalop.hip

struct CartN {
    int64_t n[3];
};

// ------------------------------------------------------------ trajectories
__global__ void int_to_float_kernel(const int32_t* __restrict__ in, int64_t n,
                                    float* __restrict__ out, int* __restrict__ bad) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = in[i];
        if (v > (1 << 24) || v < -(1 << 24)) *bad = 1;
        out[i] = (float)v;
    }
}

// An integer trajectory through the sample selection of the non-Cartesian
// path (the values are exact in Float32 below 2^24).
msr_traj* cart_traj_create(msr_ctx* ctx, int d, int64_t nsamp, int64_t nt, const int32_t* trj,
                           const uint8_t* mask, hipStream_t st) {
    MSR_REQUIRE(trj != nullptr, MSR_ERR_ARGUMENT, "trj is NULL");
    MSR_REQUIRE(d == 2 || d == 3, MSR_ERR_ARGUMENT, "trajectory dimension %d, expected 2 or 3", d);
    MSR_REQUIRE(nsamp >= 1 && nt >= 1, MSR_ERR_ARGUMENT, "empty trajectory (%lld x %lld)",
                (long long)nsamp, (long long)nt);
    const int64_t n = (int64_t)d * nsamp * nt;
    DevBuf tf, bad;
    tf.alloc(sizeof(float) * n);
    bad.alloc(sizeof(int));
    fill_floats(bad.p, 1, 0.f, st);
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(int_to_float_kernel, dim3(grid), dim3(256), 0, st, trj, n, tf.as<float>(),
                       bad.as<int>());
    MSR_CHECK_LAUNCH();
    int hbad = 0;
    MSR_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    MSR_REQUIRE(hbad == 0, MSR_ERR_ARGUMENT, "integer k-space index beyond +-2^24");
    std::unique_ptr<msr_traj> t(traj_create(ctx, d, nsamp, nt, tf.as<float>(), mask, st));
    MSR_HIP(hipStreamSynchronize(st));  // tf is released on return
    t->integer = true;
    return t.release();
}

__device__ __forceinline__ int64_t cart_cell(const float* __restrict__ kc, int64_t j, int d,
                                             const CartN& g) {
    int64_t q = 0, stride = 1;
    for (int a = 0; a < d; ++a) {
        const int64_t n = g.n[a];
        const int64_t v = (int64_t)kc[j * d + a];
        int64_t c = (v - n / 2 - 1) % n;
        if (c < 0) c += n;
        q += c * stride;
        stride *= n;
    }
    return q;
}

__global__ void cart_cells_kernel(const float* __restrict__ kc, int64_t M, int d, CartN g,
                                  int* __restrict__ cell, int* __restrict__ idx,
                                  int* __restrict__ counts) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < M;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int q = (int)cart_cell(kc, j, d, g);
        cell[j] = q;
        idx[j] = (int)j;
        atomicAdd(&counts[q], 1);  // integer counts: order-free
    }
}

// Masked samples sorted by cell (stable radix sort: a cell's samples keep
// their sample order) and the cells' offsets, per (trajectory, N).
const CartPlan& cart_plan(const msr_traj* traj, const int64_t* N, int d, hipStream_t st) {
    std::lock_guard<std::recursive_mutex> lock(traj->mu);
    std::array<int64_t, 3> key{N[0], N[1], d == 3 ? N[2] : 1};
    auto it = traj->cart_plans.find(key);
    if (it != traj->cart_plans.end()) return *it->second;
    auto p = std::make_unique<CartPlan>();
    p->G = 1;
    CartN g;
    for (int a = 0; a < 3; ++a) {
        p->n[a] = key[a];
        g.n[a] = key[a];
        p->G *= key[a];
    }
    MSR_REQUIRE(traj->M < (int64_t)INT32_MAX && p->G < (int64_t)INT32_MAX, MSR_ERR_UNSUPPORTED,
                "more than 2^31 samples or cells");
    const int M = (int)traj->M;
    DevBuf cell, cell_s, idx, counts, temp;
    cell.alloc(sizeof(int) * M);
    cell_s.alloc(sizeof(int) * M);
    idx.alloc(sizeof(int) * M);
    counts.alloc(sizeof(int) * (p->G + 1));
    p->perm.alloc(sizeof(int) * M);
    p->offs.alloc(sizeof(int) * (p->G + 1));
    MSR_HIP(hipMemsetAsync(counts.p, 0, counts.bytes, st));
    const int grid = (int)std::min<int64_t>((traj->M + 255) / 256, 8192);
    hipLaunchKernelGGL(cart_cells_kernel, dim3(grid), dim3(256), 0, st, traj->kc.as<float>(),
                       traj->M, d, g, cell.as<int>(), idx.as<int>(), counts.as<int>());
    MSR_CHECK_LAUNCH();
    int bits = 1;
    while ((int64_t(1) << bits) < p->G) ++bits;
    size_t tb1 = 0, tb2 = 0;
    MSR_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, cell.as<int>(), cell_s.as<int>(),
                                               idx.as<int>(), p->perm.as<int>(), M, 0, bits, st));
    MSR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, counts.as<int>(), p->offs.as<int>(),
                                             (int)(p->G + 1), st));
    temp.alloc(std::max<size_t>(std::max(tb1, tb2), 16));
    MSR_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, tb1, cell.as<int>(), cell_s.as<int>(),
                                               idx.as<int>(), p->perm.as<int>(), M, 0, bits, st));
    MSR_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb2, counts.as<int>(), p->offs.as<int>(),
                                             (int)(p->G + 1), st));
    MSR_HIP(hipStreamSynchronize(st));  // temporaries are freed on return
    auto& ref = *p;
    traj->cart_plans[key] = std::move(p);
    return ref;
}

// ------------------------------------------------------------------ kernel
template <bool REAL>
__device__ __forceinline__ double2 u_at(const void* U, int64_t t, int k, int64_t nt) {
    if (REAL) return make_double2((double)static_cast<const float*>(U)[t + k * nt], 0.0);
    const float2 v = static_cast<const float2*>(U)[t + k * nt];
    return make_double2(v.x, v.y);
}

// acc[p] += m * conj(u_j) u_k over the pairs j <= k (p counts them column by column)
template <int K, bool REAL>
__device__ __forceinline__ void cart_accumulate(double2 (&acc)[K * (K + 1) / 2], const void* U,
                                                int64_t t, int64_t nt, double m) {
    double2 u[K];
#pragma unroll
    for (int k = 0; k < K; ++k) u[k] = u_at<REAL>(U, t, k, nt);
    int p = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j <= k; ++j, ++p) {
            acc[p].x += m * (u[j].x * u[k].x + u[j].y * u[k].y);
            acc[p].y += m * (u[j].x * u[k].y - u[j].y * u[k].x);
        }
}

// Writes the full K x K block of cell q (lower triangle = conjugate) and its
// kmask flag: Λ[1,1] != 0 after rounding to the kernel's element type.
template <int K, bool REAL>
__device__ __forceinline__ void cart_store(const double2 (&acc)[K * (K + 1) / 2], int64_t q,
                                           void* lam, uint8_t* flag) {
    const bool keep = REAL ? ((float)acc[0].x != 0.f)
                           : ((float)acc[0].x != 0.f || (float)acc[0].y != 0.f);
    flag[q] = keep ? 1 : 0;
    int p = 0;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int j = 0; j <= k; ++j, ++p) {
            const double2 v = keep ? acc[p] : make_double2(0.0, 0.0);
            const int64_t o1 = q * K * K + j + K * k, o2 = q * K * K + k + K * j;
            if (REAL) {
                static_cast<float*>(lam)[o1] = (float)v.x;
                static_cast<float*>(lam)[o2] = (float)v.x;
            } else {
                static_cast<float2*>(lam)[o1] = make_float2((float)v.x, (float)v.y);
                static_cast<float2*>(lam)[o2] = make_float2((float)v.x, -(float)v.y);
            }
        }
}

// calculate_kernel_cartesian(img_shape, trj, U; sample_mask) (src/FFTNormalOp.jl:81-99)
template <int K, bool REAL>
__global__ __launch_bounds__(256) void cart_kernel_traj(const int* __restrict__ perm,
                                                        const int* __restrict__ offs,
                                                        const int* __restrict__ frame,
                                                        const void* __restrict__ U, int64_t nt,
                                                        int64_t G, void* __restrict__ lam,
                                                        uint8_t* __restrict__ flag) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < G;
         q += (int64_t)gridDim.x * blockDim.x) {
        double2 acc[K * (K + 1) / 2];
#pragma unroll
        for (int p = 0; p < K * (K + 1) / 2; ++p) acc[p] = make_double2(0.0, 0.0);
        for (int e = offs[q]; e < offs[q + 1]; ++e)
            cart_accumulate<K, REAL>(acc, U, frame[perm[e]], nt, 1.0);
        cart_store<K, REAL>(acc, q, lam, flag);
    }
}

// calculate_kernel_cartesian(M, U) (src/FFTNormalOp.jl:101-111): M (N..., Nt)
// Bool, ifftshifted over the image axes: cell q reads position (q + N/2) mod N.
template <int K, bool REAL>
__global__ __launch_bounds__(256) void cart_kernel_mask(const uint8_t* __restrict__ M, CartN g,
                                                        int64_t nt, const void* __restrict__ U,
                                                        int64_t G, void* __restrict__ lam,
                                                        uint8_t* __restrict__ flag) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < G;
         q += (int64_t)gridDim.x * blockDim.x) {
        int64_t src = 0, stride = 1, r = q;
        for (int a = 0; a < 3; ++a) {
            const int64_t c = r % g.n[a];
            r /= g.n[a];
            src += ((c + g.n[a] / 2) % g.n[a]) * stride;
            stride *= g.n[a];
        }
        double2 acc[K * (K + 1) / 2];
#pragma unroll
        for (int p = 0; p < K * (K + 1) / 2; ++p) acc[p] = make_double2(0.0, 0.0);
        for (int64_t t = 0; t < nt; ++t)
            if (M[src + t * G]) cart_accumulate<K, REAL>(acc, U, t, nt, 1.0);
        cart_store<K, REAL>(acc, q, lam, flag);
    }
}

// FFTNormalOp(Λ) (src/FFTNormalOp.jl:38-47): the given (K, K, N...) kernel,
// zero on the cells off kmask = Λ[1,1] != 0 (kept in full, as the reference).
template <class LT>
__global__ void cart_kernel_dense(const LT* __restrict__ in, int KK, int64_t G,
                                  LT* __restrict__ lam, uint8_t* __restrict__ flag) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < G * KK;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = e / KK;
        const LT l11 = in[q * KK];
        bool keep;
        if constexpr (sizeof(LT) == sizeof(float)) keep = l11 != 0.f;
        else keep = l11.x != 0.f || l11.y != 0.f;
        if (e == q * KK) flag[q] = keep ? 1 : 0;
        lam[e] = keep ? in[e] : LT{};
    }
}

__global__ void flag_counts_kernel(const uint8_t* __restrict__ flag, int64_t G,
                                   int* __restrict__ c) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < G;
         q += (int64_t)gridDim.x * blockDim.x)
        c[q] = flag[q];
}

__global__ void kmask_scatter_kernel(const uint8_t* __restrict__ flag, const int* __restrict__ pos,
                                     int64_t G, int64_t* __restrict__ kmask) {
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < G;
         q += (int64_t)gridDim.x * blockDim.x)
        if (flag[q]) kmask[pos[q]] = q + 1;
}

// kmask (1-based, ascending) from the per-cell flags
static void cart_kmask(msr_normalop* op, const uint8_t* flag, int64_t G, hipStream_t st) {
    DevBuf c, pos, temp;
    c.alloc(sizeof(int) * (G + 1));
    pos.alloc(sizeof(int) * (G + 1));
    MSR_HIP(hipMemsetAsync(c.p, 0, c.bytes, st));
    const int grid = (int)std::min<int64_t>((G + 255) / 256, 8192);
    hipLaunchKernelGGL(flag_counts_kernel, dim3(grid), dim3(256), 0, st, flag, G, c.as<int>());
    MSR_CHECK_LAUNCH();
    size_t tb = 0;
    MSR_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, c.as<int>(), pos.as<int>(), (int)(G + 1), st));
    temp.alloc(std::max<size_t>(tb, 16));
    MSR_HIP(hipcub::DeviceScan::ExclusiveSum(temp.p, tb, c.as<int>(), pos.as<int>(), (int)(G + 1), st));
    int nkm = 0;
    MSR_HIP(hipMemcpyAsync(&nkm, pos.as<int>() + G, sizeof(int), hipMemcpyDeviceToHost, st));
    MSR_HIP(hipStreamSynchronize(st));
    op->nkm = nkm;
    op->kmask->alloc(sizeof(int64_t) * std::max<int64_t>(nkm, 1));
    hipLaunchKernelGGL(kmask_scatter_kernel, dim3(grid), dim3(256), 0, st, flag, pos.as<int>(), G,
                       op->kmask->as<int64_t>());
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipStreamSynchronize(st));
}

#define MSR_CART_K(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8)

// Kernel of a Cartesian operator from a trajectory (mask == nullptr) or a
// sampling mask M (traj == nullptr).
void cart_build_kernel(msr_normalop* op, const msr_traj* traj, const uint8_t* M, int64_t nt_mask,
                       const void* U, int u_real, hipStream_t st) {
    const int64_t G = op->V;
    const int K = op->K;
    CartN g{{op->N[0], op->N[1], op->N[2]}};
    const size_t es = u_real ? sizeof(float) : sizeof(float2);
    op->lam->alloc(es * (size_t)K * K * G);
    DevBuf flag;
    flag.alloc(G);
    const int grid = (int)std::min<int64_t>((G + 255) / 256, 16384);
    if (traj) {
        MSR_REQUIRE(traj->integer, MSR_ERR_ARGUMENT,
                    "FFTNormalOp needs an integer (Cartesian) trajectory");
        MSR_REQUIRE(traj->d == op->d, MSR_ERR_ARGUMENT, "trajectory has %d dimensions, img_shape %d",
                    traj->d, op->d);
        const CartPlan& plan = cart_plan(traj, op->N, op->d, st);
#define MSR_CART_T(KK)                                                                           \
    if (K == KK) {                                                                               \
        if (u_real)                                                                              \
            hipLaunchKernelGGL((cart_kernel_traj<KK, true>), dim3(grid), dim3(256), 0, st,       \
                               plan.perm.as<int>(), plan.offs.as<int>(), traj->frame.as<int>(), \
                               U, traj->nt, G, op->lam->p, flag.as<uint8_t>());                  \
        else                                                                                     \
            hipLaunchKernelGGL((cart_kernel_traj<KK, false>), dim3(grid), dim3(256), 0, st,      \
                               plan.perm.as<int>(), plan.offs.as<int>(), traj->frame.as<int>(), \
                               U, traj->nt, G, op->lam->p, flag.as<uint8_t>());                  \
    }
        MSR_CART_K(MSR_CART_T)
#undef MSR_CART_T
    } else {
#define MSR_CART_M(KK)                                                                          \
    if (K == KK) {                                                                              \
        if (u_real)                                                                             \
            hipLaunchKernelGGL((cart_kernel_mask<KK, true>), dim3(grid), dim3(256), 0, st, M, g, \
                               nt_mask, U, G, op->lam->p, flag.as<uint8_t>());                  \
        else                                                                                    \
            hipLaunchKernelGGL((cart_kernel_mask<KK, false>), dim3(grid), dim3(256), 0, st, M,  \
                               g, nt_mask, U, G, op->lam->p, flag.as<uint8_t>());               \
    }
        MSR_CART_K(MSR_CART_M)
#undef MSR_CART_M
    }
    MSR_CHECK_LAUNCH();
    cart_kmask(op, flag.as<uint8_t>(), G, st);
}

void cart_load_dense(msr_normalop* op, const void* lambda, hipStream_t st) {
    const int64_t G = op->V;
    const int KK = op->K * op->K;
    const size_t es = op->lam_real ? sizeof(float) : sizeof(float2);
    op->lam->alloc(es * (size_t)KK * G);
    DevBuf flag;
    flag.alloc(G);
    const int grid = (int)std::min<int64_t>((G * KK + 255) / 256, 16384);
    if (op->lam_real)
        hipLaunchKernelGGL(cart_kernel_dense<float>, dim3(grid), dim3(256), 0, st,
                           static_cast<const float*>(lambda), KK, G, op->lam->as<float>(),
                           flag.as<uint8_t>());
    else
        hipLaunchKernelGGL(cart_kernel_dense<float2>, dim3(grid), dim3(256), 0, st,
                           static_cast<const float2*>(lambda), KK, G, op->lam->as<float2>(),
                           flag.as<uint8_t>());
    MSR_CHECK_LAUNCH();
    cart_kmask(op, flag.as<uint8_t>(), G, st);
}

// ------------------------------------------------------------------ apply
template <class AT>
__global__ __launch_bounds__(256) void cart_expand_kernel(const AT* __restrict__ b,
                                                          const float2* __restrict__ cmaps, int C,
                                                          int K, int64_t V, AT* __restrict__ T,
                                                          const int* __restrict__ skip) {
    if (skip && *skip) return;
    using RT = real_t<AT>;
    const int64_t n = (int64_t)C * K * V;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i % V;
        const int64_t ck = i / V;
        const int k = (int)(ck % K);
        const int c = (int)(ck / K);
        AT x = b[(int64_t)k * V + v];
        if (cmaps) {
            const float2 s = cmaps[(int64_t)c * V + v];
            x = cmul(AT{(RT)s.x, (RT)s.y}, x);
        }
        T[i] = x;
    }
}

// kL2[j] = sum_k Λ[j, k] kL1[k] on every cell (zero off the kmask), each coil
// (src/FFTNormalOp.jl:133-135); one thread per cell reads its K x K block once.
template <int K, class AT, class LT>
__global__ __launch_bounds__(256) void cart_multiply_kernel(const AT* __restrict__ in,
                                                            AT* __restrict__ out,
                                                            const LT* __restrict__ lam, int C,
                                                            int64_t G,
                                                            const int* __restrict__ skip) {
    if (skip && *skip) return;
    using RT = real_t<AT>;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < G;
         q += (int64_t)gridDim.x * blockDim.x) {
        LT L[K * K];
#pragma unroll
        for (int e = 0; e < K * K; ++e) L[e] = lam[q * K * K + e];
        for (int c = 0; c < C; ++c) {
            AT x[K];
#pragma unroll
            for (int k = 0; k < K; ++k) x[k] = in[((int64_t)c * K + k) * G + q];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                AT acc{0, 0};
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if constexpr (sizeof(LT) == sizeof(float)) {
                        acc.x += (RT)L[j + K * k] * x[k].x;
                        acc.y += (RT)L[j + K * k] * x[k].y;
                    } else {
                        const AT l{(RT)L[j + K * k].x, (RT)L[j + K * k].y};
                        acc = cadd(acc, cmul(l, x[k]));
                    }
                }
                out[((int64_t)c * K + j) * G + q] = acc;
            }
        }
    }
}

template <class AT>
__global__ __launch_bounds__(256) void cart_combine_kernel(const AT* __restrict__ T,
                                                           const float2* __restrict__ cmaps,
                                                           int C, int K, int64_t V,
                                                           AT* __restrict__ partial,
                                                           const int* __restrict__ skip) {
    if (skip && *skip) return;
    using RT = real_t<AT>;
    const int64_t n = (int64_t)K * V;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = i % V;
        const int k = (int)(i / V);
        AT acc{0, 0};
        for (int c = 0; c < C; ++c) {
            const AT x = T[((int64_t)c * K + k) * V + v];
            if (cmaps) {
                const float2 s = cmaps[(int64_t)c * V + v];
                acc = cadd(acc, cmulc(AT{(RT)s.x, (RT)s.y}, x));
            } else {
                acc = cadd(acc, x);
            }
        }
        partial[i] = acc;
    }
}

// FFT along every image axis of nch grids (in -> out); the y and z passes
// ping-pong through `mid`, the result lands in out.  dir -1 forward, +1
// inverse (unnormalised); sel: output placement (SEL_FIRST or SEL_CENTER).
static void cart_fft(msr_ctx* ctx, int d, const int64_t* N, int64_t nch, const void* in,
                     void* out, void* mid, int dir, int sel, bool f64, const int* skip,
                     hipStream_t st) {
    const int64_t nx = N[0], ny = N[1], nz = d == 3 ? N[2] : 1, V = nx * ny * nz;
    auto line = [&](int a) {
        LinePass p;
        p.f64 = f64 ? 1 : 0;
        p.nf = (int)N[a];
        p.dir = dir;
        p.lin = p.lout = (int)N[a];
        p.isel = SEL_FIRST;
        p.sel = sel;
        p.skip = skip;
        return p;
    };
    // 2D: in -x-> mid -y-> out;  3D: in -x-> out -y-> mid -z-> out
    LinePass px = line(0);
    px.in = in;
    px.out = d == 3 ? out : mid;
    px.nlines = nch * ny * nz;
    px.in_pitch = px.out_pitch = nx;
    run_rows(px, ctx, st);
    LinePass py = line(1);
    py.in = px.out;
    py.out = d == 3 ? mid : out;
    py.nb = nch * nz;
    py.ncols = nx;
    py.in_bstride = py.out_bstride = ny * nx;
    py.in_stride = py.out_stride = nx;
    run_cols(py, ctx, st);
    if (d == 3) {
        LinePass pz = line(2);
        pz.in = mid;
        pz.out = out;
        pz.nb = nch;
        pz.ncols = ny * nx;
        pz.in_bstride = pz.out_bstride = V;
        pz.in_stride = pz.out_stride = ny * nx;
        run_cols(pz, ctx, st);
    }
}

template <class AT>
static void cart_passes_t(msr_normalop* op, const AT* b, hipStream_t st, const int* skip) {
    const int C = op->C, K = op->K;
    const int64_t V = op->V;
    const float2* cm = op->identity_cmaps ? nullptr : op->cmaps.as<float2>();
    AT* A = op->T.as<AT>();
    AT* B = op->T2.as<AT>();
    AT* P = op->partial.as<AT>();
    const bool f64 = sizeof(AT) == sizeof(double2);
    const int64_t n = (int64_t)C * K * V;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(cart_expand_kernel<AT>, dim3(grid), dim3(256), 0, st, b, cm, C, K, V, A, skip);
    MSR_CHECK_LAUNCH();
    // forward FFT A -> B (A also holds the middle stage: the x pass consumes it first)
    cart_fft(op->ctx, op->d, op->N, (int64_t)C * K, A, B, A, -1, SEL_FIRST, f64, skip, st);
    const int gq = (int)std::min<int64_t>((V + 255) / 256, 16384);
#define MSR_CART_MUL(KK)                                                                         \
    if (K == KK) {                                                                               \
        if (op->lam_real)                                                                        \
            hipLaunchKernelGGL((cart_multiply_kernel<KK, AT, float>), dim3(gq), dim3(256), 0, st, \
                               B, A, op->lam->as<float>(), C, V, skip);                          \
        else                                                                                     \
            hipLaunchKernelGGL((cart_multiply_kernel<KK, AT, float2>), dim3(gq), dim3(256), 0,   \
                               st, B, A, op->lam->as<float2>(), C, V, skip);                     \
    }
    MSR_CART_K(MSR_CART_MUL)
#undef MSR_CART_MUL
    MSR_CHECK_LAUNCH();
    cart_fft(op->ctx, op->d, op->N, (int64_t)C * K, A, B, A, +1, SEL_FIRST, f64, skip, st);
    const int gk = (int)std::min<int64_t>(((int64_t)K * V + 255) / 256, 16384);
    hipLaunchKernelGGL(cart_combine_kernel<AT>, dim3(gk), dim3(256), 0, st, B, cm, C, K, V, P, skip);
    MSR_CHECK_LAUNCH();
}

void launch_cart_passes(msr_normalop* op, const void* b, bool f64, hipStream_t st,
                        const int* skip) {
    if (f64) cart_passes_t<double2>(op, static_cast<const double2*>(b), st, skip);
    else cart_passes_t<float2>(op, static_cast<const float2*>(b), st, skip);
}

// ------------------------------------------------------------ backprojection
// dataU[cell, c, k] = sum over the cell's samples (sample order, ComplexF32 as
// the reference) of data * conj(U[t, k])  (src/BackProjection.jl:113-121).
// One thread per (cell, coil) holds CART_KB coefficients in registers, so a
// sample is read once per CART_KB coefficients; neighbouring threads are
// neighbouring readout cells, whose e-th samples are neighbouring samples of
// one readout line (coalesced data reads).
constexpr int CART_KB = 8;
__global__ __launch_bounds__(256) void cart_grid_kernel(
    const int* __restrict__ perm, const int* __restrict__ offs, const int* __restrict__ frame,
    const int64_t* __restrict__ sel, const float2* __restrict__ data, int64_t nsel,
    const void* __restrict__ U, int u_real, int64_t nt, int K, int C, int64_t G,
    float2* __restrict__ grid) {
    const int64_t n = G * C;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = i % G;
        const int c = (int)(i / G);
        const int e0 = offs[q], e1 = offs[q + 1];
        const float2* yc = data + (int64_t)c * nsel;
        for (int k0 = 0; k0 < K; k0 += CART_KB) {
            float2 acc[CART_KB];
#pragma unroll
            for (int kk = 0; kk < CART_KB; ++kk) acc[kk] = make_float2(0.f, 0.f);
            auto add = [&](int t, float2 y) {
#pragma unroll
                for (int kk = 0; kk < CART_KB; ++kk) {
                    const int k = k0 + kk;
                    if (k >= K) break;
                    float2 u;
                    if (!U) u = make_float2(t == k ? 1.f : 0.f, 0.f);
                    else if (u_real)
                        u = make_float2(static_cast<const float*>(U)[t + (int64_t)k * nt], 0.f);
                    else u = static_cast<const float2*>(U)[t + (int64_t)k * nt];
                    acc[kk] = cadd(acc[kk], cmulc(u, y));  // y * conj(u)
                }
            };
            // four samples' loads issued together, then added in sample order
            int e = e0;
            for (; e + 4 <= e1; e += 4) {
                int j[4], t[4];
                float2 y[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) j[r] = perm[e + r];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    t[r] = frame[j[r]];
                    y[r] = yc[sel ? sel[j[r]] : j[r]];
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) add(t[r], y[r]);
            }
            for (; e < e1; ++e) {
                const int jj = perm[e];
                add(frame[jj], yc[sel ? sel[jj] : jj]);
            }
#pragma unroll
            for (int kk = 0; kk < CART_KB; ++kk)
                if (k0 + kk < K) grid[((int64_t)c * K + k0 + kk) * G + q] = acc[kk];
        }
    }
}

// out = scale * sum_c conj(S_c) img_ck (combined) or scale * img (per coil)
__global__ __launch_bounds__(256) void cart_bp_out_kernel(const float2* __restrict__ img,
                                                          const float2* const* __restrict__ cm,
                                                          int C, int K, int64_t V, float scale,
                                                          float2* __restrict__ out) {
    const int64_t n = cm ? (int64_t)K * V : (int64_t)C * K * V;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (!cm) {
            out[i] = cscale(img[i], scale);
            continue;
        }
        const int64_t v = i % V;
        const int k = (int)(i / V);
        float2 acc = make_float2(0.f, 0.f);
        for (int c = 0; c < C; ++c)
            acc = cadd(acc, cmulc(cm[c][v], cscale(img[((int64_t)c * K + k) * V + v], scale)));
        out[i] = acc;
    }
}

// calculate_backprojection(data, trj::Integer, cmaps | img_shape; U, sample_mask)
// (src/BackProjection.jl:104-156): grid, normalised inverse FFT, fftshift,
// conj(S_c) coil sum or per-coil images (C, K, N...).
void cart_backprojection(msr_ctx* ctx, const msr_traj* traj, int d, const int64_t* N,
                         const void* data, int ncoils, int K, const void* U, int u_real,
                         const void* const* cmaps, void* out, hipStream_t st) {
    MSR_REQUIRE(traj->d == d, MSR_ERR_ARGUMENT, "trajectory has %d dimensions, img_shape %d",
                traj->d, d);
    int64_t V = 1;
    for (int a = 0; a < d; ++a) {
        MSR_REQUIRE(N[a] >= 1 && fft_len_supported(N[a]), MSR_ERR_UNSUPPORTED,
                    "Cartesian grid size N[%d] = %lld has no FFT kernel in this build", a,
                    (long long)N[a]);
        V *= N[a];
    }
    const CartPlan& plan = cart_plan(traj, N, d, st);
    const int64_t nch = (int64_t)ncoils * K;
    DevBuf work;
    work.alloc(2 * sizeof(float2) * (size_t)nch * V);
    float2* Gd = work.as<float2>();
    float2* H = Gd + nch * V;
    const int grid = (int)std::min<int64_t>(((int64_t)ncoils * V + 255) / 256, 16384);
    hipLaunchKernelGGL(cart_grid_kernel, dim3(grid), dim3(256), 0, st, plan.perm.as<int>(),
                       plan.offs.as<int>(), traj->frame.as<int>(),
                       traj->sel.p ? traj->sel.as<int64_t>() : nullptr, (const float2*)data,
                       traj->nsamp * traj->nt, U, u_real, traj->nt, K, ncoils, V, Gd);
    MSR_CHECK_LAUNCH();
    // inverse FFT with fftshifted output: Gd -> H (3D through Gd)
    cart_fft(ctx, d, N, nch, Gd, H, Gd, +1, SEL_CENTER, false, nullptr, st);
    // the host staging array lives until the stream synchronisation below
    DevBuf ptrs;
    std::vector<const float2*> hp;
    if (cmaps) {
        hp.resize(ncoils);
        for (int c = 0; c < ncoils; ++c) {
            MSR_REQUIRE(cmaps[c] != nullptr, MSR_ERR_ARGUMENT, "cmaps[%d] is NULL", c);
            hp[c] = (const float2*)cmaps[c];
        }
        ptrs.alloc(sizeof(float2*) * ncoils);
        MSR_HIP(hipMemcpyAsync(ptrs.p, hp.data(), sizeof(float2*) * ncoils, hipMemcpyHostToDevice, st));
    }
    const int64_t nout = cmaps ? (int64_t)K * V : nch * V;
    const int go = (int)std::min<int64_t>((nout + 255) / 256, 16384);
    hipLaunchKernelGGL(cart_bp_out_kernel, dim3(go), dim3(256), 0, st, H,
                       cmaps ? ptrs.as<const float2*>() : nullptr, ncoils, K, V,
                       (float)(1.0 / (double)V), (float2*)out);
    MSR_CHECK_LAUNCH();
    MSR_HIP(hipStreamSynchronize(st));  // the workspace is released on return
}

}  // namespace msr
