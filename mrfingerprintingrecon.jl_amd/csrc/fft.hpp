// LDS-staged mixed-radix Stockham FFT for gfx950.
//
// A group of T threads transforms one line of N complex values held in LDS.
// Each stage of radix R reads R values per butterfly (stride N/R), applies
// the twiddles W_{NS*R}^{k r} (k = j mod NS, NS = product of earlier radices),
// runs an R-point DFT in registers and writes the R results to positions
// (j - k) R + k + r NS, so the output is in natural order without a
// bit-reversal pass (Stockham autosort).  Stages are separated by workgroup
// barriers; every thread of the workgroup must call the transform.
//
// Sign: DIR = -1 forward (exp(-2 pi i ...)), DIR = +1 inverse, unnormalised.
// Twiddles come from a table tw[m] = exp(-2 pi i m / N) computed on the host in
// double precision (stored as ComplexF32 or ComplexF64 for the arithmetic type)
// and staged into LDS by the caller.  The arithmetic type is the element type
// of the line: float2 or double2.
#pragma once

#include "common.hpp"

namespace msr {

// ----------------------------------------------------------- radix plans
// R[]: radices in stage order; T: threads per line (divides N/R for the
// stages where that matters; other stages use a guarded loop).
template <int N> struct FFTPlan;
#define MSR_PLAN(NN, TT, ...)                                   \
    template <> struct FFTPlan<NN> {                            \
        static constexpr int T = TT;                            \
        static constexpr int R[] = {__VA_ARGS__};               \
        static constexpr int NST = sizeof(R) / sizeof(int);     \
    };
MSR_PLAN(16, 4, 4, 4)
MSR_PLAN(24, 1, 8, 3)
MSR_PLAN(32, 4, 8, 4)
MSR_PLAN(40, 1, 8, 5)
MSR_PLAN(48, 2, 8, 3, 2)
MSR_PLAN(64, 8, 8, 8)
MSR_PLAN(80, 2, 8, 5, 2)
MSR_PLAN(96, 4, 8, 4, 3)
MSR_PLAN(128, 16, 8, 4, 4)
MSR_PLAN(160, 8, 8, 4, 5)
MSR_PLAN(192, 8, 8, 8, 3)
MSR_PLAN(256, 32, 8, 8, 4)
MSR_PLAN(320, 16, 8, 8, 5)
MSR_PLAN(384, 16, 8, 8, 3, 2)
MSR_PLAN(512, 64, 8, 8, 8)
MSR_PLAN(640, 16, 8, 8, 5, 2)
MSR_PLAN(768, 32, 8, 8, 4, 3)
MSR_PLAN(1024, 64, 8, 8, 4, 4)
#undef MSR_PLAN

// Supported line lengths (each has a plan above).
#define MSR_FFT_LENGTHS(X) \
    X(16) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(320) X(384) \
    X(512) X(640) X(768) X(1024)

// Invoke F.template operator()<N>() for a runtime length n; returns false if
// n has no plan.
template <class F>
inline bool dispatch_fft_len(int n, F&& f) {
#define MSR_CASE(NN)                     \
    case NN:                             \
        f.template operator()<NN>();     \
        return true;
    switch (n) {
        MSR_FFT_LENGTHS(MSR_CASE)
        default: return false;
    }
#undef MSR_CASE
}

inline bool fft_len_supported(int64_t n) {
#define MSR_CASE(NN) \
    case NN: return true;
    switch (n) {
        MSR_FFT_LENGTHS(MSR_CASE)
        default: return false;
    }
#undef MSR_CASE
}

// ----------------------------------------------------------- butterflies
// multiply by DIR * i
template <int DIR, class CT>
__device__ __forceinline__ CT mul_i(CT a) {
    CT r;
    if (DIR > 0) {
        r.x = -a.y;
        r.y = a.x;
    } else {
        r.x = a.y;
        r.y = -a.x;
    }
    return r;
}

// Butterflies of radix R on values of complex type CT (float2 or double2);
// constants are rounded to the arithmetic type.
template <int R, int DIR> struct Butterfly;

template <int DIR> struct Butterfly<2, DIR> {
    template <class CT>
    __device__ __forceinline__ static void run(CT* v) {
        CT a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR> struct Butterfly<4, DIR> {
    template <class CT>
    __device__ __forceinline__ static void run(CT* v) {
        CT t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        CT t2 = cadd(v[1], v[3]), t3 = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int DIR> struct Butterfly<8, DIR> {
    template <class CT>
    __device__ __forceinline__ static void run(CT* v) {
        using RT = real_t<CT>;
        constexpr RT s = (RT)0.70710678118654752440;
        CT e[4] = {v[0], v[2], v[4], v[6]};
        CT o[4] = {v[1], v[3], v[5], v[7]};
        Butterfly<4, DIR>::run(e);
        Butterfly<4, DIR>::run(o);
        // w8^q for q = 0..3 with w8 = (1 + DIR i)/sqrt2
        CT o1 = cscale(cadd(o[1], mul_i<DIR>(o[1])), s);
        CT o2 = mul_i<DIR>(o[2]);
        CT o3 = cscale(csub(mul_i<DIR>(o[3]), o[3]), s);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

template <int DIR> struct Butterfly<3, DIR> {
    template <class CT>
    __device__ __forceinline__ static void run(CT* v) {
        using RT = real_t<CT>;
        constexpr RT h = (RT)0.86602540378443864676;  // sin(2 pi / 3)
        CT s = cadd(v[1], v[2]);
        CT d = mul_i<DIR>(cscale(csub(v[1], v[2]), h));
        CT m;
        m.x = v[0].x - (RT)0.5 * s.x;
        m.y = v[0].y - (RT)0.5 * s.y;
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, d);
        v[2] = csub(m, d);
    }
};

template <int DIR> struct Butterfly<5, DIR> {
    template <class CT>
    __device__ __forceinline__ static void run(CT* v) {
        using RT = real_t<CT>;
        constexpr RT c1 = (RT)0.30901699437494742410;   // cos(2 pi/5)
        constexpr RT c2 = (RT)-0.80901699437494742410;  // cos(4 pi/5)
        constexpr RT s1 = (RT)0.95105651629515357212;   // sin(2 pi/5)
        constexpr RT s2 = (RT)0.58778525229247312917;   // sin(4 pi/5)
        CT a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        CT a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        CT p1, p2, q1, q2;
        p1.x = v[0].x + c1 * a1.x + c2 * a2.x;
        p1.y = v[0].y + c1 * a1.y + c2 * a2.y;
        p2.x = v[0].x + c2 * a1.x + c1 * a2.x;
        p2.y = v[0].y + c2 * a1.y + c1 * a2.y;
        q1.x = s1 * b1.x + s2 * b2.x;
        q1.y = s1 * b1.y + s2 * b2.y;
        q2.x = s2 * b1.x - s1 * b2.x;
        q2.y = s2 * b1.y - s1 * b2.y;
        q1 = mul_i<DIR>(q1);
        q2 = mul_i<DIR>(q2);
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(p1, q1);
        v[4] = csub(p1, q1);
        v[2] = cadd(p2, q2);
        v[3] = csub(p2, q2);
    }
};

// ----------------------------------------------------------- LDS layout
// Element i of a line lives at lds_pad(i): one complex of padding after every
// 8 elements, so the stride-R stores of the first Stockham stages and the
// strided reads of the transposing passes spread over the 64 LDS banks.
__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 3); }
template <int N>
constexpr int lds_line() { return N + N / 8; }

// ----------------------------------------------------------- stages
template <int N, int R, int NS, int T, int DIR, class CT>
__device__ __forceinline__ void fft_stage(CT* buf, const CT* tw, int t, bool active) {
    constexpr int NB = N / R;
    constexpr int PER = (NB + T - 1) / T;
    constexpr bool EXACT = (NB % T) == 0;
    CT v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * T;
        if (active && (EXACT || j < NB)) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = buf[lds_pad(j + r * NB)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * T;
        if (active && (EXACT || j < NB)) {
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    CT w = tw[(k * r) * (N / (NS * R))];
                    if constexpr (DIR > 0) w.y = -w.y;
                    v[p][r] = cmul(v[p][r], w);
                }
            }
            Butterfly<R, DIR>::run(v[p]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[lds_pad(base + r * NS)] = v[p][r];
        }
    }
    __syncthreads();
}

template <int N, int T, int DIR, int S, int NS, class CT>
__device__ __forceinline__ void fft_stages(CT* buf, const CT* tw, int t, bool active) {
    if constexpr (S < FFTPlan<N>::NST) {
        constexpr int R = FFTPlan<N>::R[S];
        fft_stage<N, R, NS, T, DIR, CT>(buf, tw, t, active);
        fft_stages<N, T, DIR, S + 1, NS * R, CT>(buf, tw, t, active);
    }
}

// In-place N-point FFT of a line (element i at buf[lds_pad(i)]) by the
// T = FFTPlan<N>::T threads with local index t.  Requires a barrier before the
// call (data in LDS); ends with a barrier.  All threads of the workgroup must
// call it; threads with active = false only take part in the barriers (buf is
// not touched), for workgroups with more threads than line work.
template <int N, int DIR, class CT>
__device__ __forceinline__ void fft_lds(CT* buf, const CT* tw, int t, bool active = true) {
    fft_stages<N, FFTPlan<N>::T, DIR, 0, 1, CT>(buf, tw, t, active);
}

// Copy an N-entry twiddle table from global memory into LDS (no barrier).
template <int N, class CT>
__device__ __forceinline__ void load_twiddles(CT* s_tw, const CT* g_tw) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) s_tw[i] = g_tw[i];
}

}  // namespace msr
