// LDS-staged mixed-radix Stockham FFT for gfx950.
//
// A group of T threads transforms one line of N complex values held in LDS.
// Each stage of radix R reads R values per butterfly (stride N/R), applies
// the twiddles W_{NS*R}^{k r} (k = j mod NS, NS = product of earlier radices),
// runs an R-point DFT in registers and writes the R results to positions
// (j - k) R + k + r NS, so the output is in natural order without a
// bit-reversal pass (Stockham autosort).  A line's threads lie in one
// wavefront, so stages are ordered by wavefront-level fences; the first
// stage may take its inputs from registers and the last leave its outputs
// there (fft_wave), saving the LDS round trips next to global memory.
//
// Sign: DIR = -1 forward (exp(-2 pi i ...)), DIR = +1 inverse, unnormalised.
// Twiddles come from a stage-ordered table (entry NS - R0 + (r - 1) NS + k =
// exp(-2 pi i k r / (NS R)) for the stage with radix R after NS) computed on
// the host in double precision (stored as ComplexF32 or ComplexF64 for the
// arithmetic type) and staged into LDS by the caller.  The arithmetic type is the element type
// of the line: float2 or double2.
#pragma once

#include "common.hpp"

namespace msr {

// ----------------------------------------------------------- radix plans
// R[]: radices in stage order; T: threads per line, a power of two dividing 64
// (a line lies in one wavefront).  In an exact plan T also divides N/R of
// every stage (every thread does the same number of butterflies, so a thread
// holds N/T values of its line), which the register-resident transforms of
// the fused apply need; the other plans (lengths with several factors 3 or 5)
// serve the LDS transforms, whose stages skip the missing butterflies.  Radix
// choice: as many 8s as the factor 2^a allows, then a 4 or a 2, then the 5s and
// 3s; T is the largest power of two up to 64 and up to the smallest stage's
// butterfly count.
template <int N> struct FFTPlan;
#define MSR_PLAN(NN, TT, ...)                                   \
    template <> struct FFTPlan<NN> {                            \
        static constexpr int T = TT;                            \
        static constexpr int R[] = {__VA_ARGS__};               \
        static constexpr int NST = sizeof(R) / sizeof(int);     \
    };
MSR_PLAN(2, 1, 2)
MSR_PLAN(3, 1, 3)
MSR_PLAN(4, 1, 4)
MSR_PLAN(5, 1, 5)
MSR_PLAN(6, 2, 2, 3)
MSR_PLAN(8, 1, 8)
MSR_PLAN(9, 2, 3, 3)
MSR_PLAN(10, 2, 2, 5)
MSR_PLAN(12, 2, 4, 3)
MSR_PLAN(15, 2, 5, 3)
MSR_PLAN(16, 4, 4, 4)
MSR_PLAN(18, 4, 2, 3, 3)
MSR_PLAN(20, 4, 4, 5)
MSR_PLAN(24, 1, 8, 3)
MSR_PLAN(25, 4, 5, 5)
MSR_PLAN(27, 8, 3, 3, 3)
MSR_PLAN(30, 4, 2, 5, 3)
MSR_PLAN(32, 4, 8, 4)
MSR_PLAN(36, 8, 4, 3, 3)
MSR_PLAN(40, 1, 8, 5)
MSR_PLAN(45, 8, 5, 3, 3)
MSR_PLAN(48, 2, 8, 3, 2)
MSR_PLAN(50, 8, 2, 5, 5)
MSR_PLAN(54, 16, 2, 3, 3, 3)
MSR_PLAN(60, 8, 4, 5, 3)
MSR_PLAN(64, 8, 8, 8)
MSR_PLAN(72, 8, 8, 3, 3)
MSR_PLAN(75, 8, 5, 5, 3)
MSR_PLAN(80, 2, 8, 5, 2)
MSR_PLAN(81, 16, 3, 3, 3, 3)
MSR_PLAN(90, 16, 2, 5, 3, 3)
MSR_PLAN(96, 4, 8, 4, 3)
MSR_PLAN(100, 16, 4, 5, 5)
MSR_PLAN(108, 16, 4, 3, 3, 3)
MSR_PLAN(120, 8, 8, 5, 3)
MSR_PLAN(125, 16, 5, 5, 5)
MSR_PLAN(128, 16, 8, 4, 4)
MSR_PLAN(135, 16, 5, 3, 3, 3)
MSR_PLAN(144, 16, 8, 2, 3, 3)
MSR_PLAN(150, 16, 2, 5, 5, 3)
MSR_PLAN(160, 8, 4, 4, 5, 2)
MSR_PLAN(162, 32, 2, 3, 3, 3, 3)
MSR_PLAN(180, 32, 4, 5, 3, 3)
MSR_PLAN(192, 8, 8, 8, 3)
MSR_PLAN(200, 16, 8, 5, 5)
MSR_PLAN(216, 16, 8, 3, 3, 3)
MSR_PLAN(225, 32, 5, 5, 3, 3)
MSR_PLAN(240, 16, 8, 2, 5, 3)
MSR_PLAN(243, 64, 3, 3, 3, 3, 3)
MSR_PLAN(250, 32, 2, 5, 5, 5)
MSR_PLAN(256, 32, 8, 8, 4)
MSR_PLAN(270, 32, 2, 5, 3, 3, 3)
MSR_PLAN(288, 32, 8, 4, 3, 3)
MSR_PLAN(300, 32, 4, 5, 5, 3)
MSR_PLAN(320, 16, 4, 4, 4, 5)
MSR_PLAN(324, 64, 4, 3, 3, 3, 3)
MSR_PLAN(360, 32, 8, 5, 3, 3)
MSR_PLAN(375, 64, 5, 5, 5, 3)
MSR_PLAN(384, 16, 8, 8, 3, 2)
MSR_PLAN(400, 32, 8, 2, 5, 5)
MSR_PLAN(405, 64, 5, 3, 3, 3, 3)
MSR_PLAN(432, 32, 8, 2, 3, 3, 3)
MSR_PLAN(450, 64, 2, 5, 5, 3, 3)
MSR_PLAN(480, 32, 8, 4, 5, 3)
MSR_PLAN(486, 64, 2, 3, 3, 3, 3, 3)
MSR_PLAN(500, 64, 4, 5, 5, 5)
MSR_PLAN(512, 64, 8, 8, 8)
MSR_PLAN(540, 64, 4, 5, 3, 3, 3)
MSR_PLAN(576, 64, 8, 8, 3, 3)
MSR_PLAN(600, 64, 8, 5, 5, 3)
MSR_PLAN(625, 64, 5, 5, 5, 5)
MSR_PLAN(640, 16, 8, 8, 5, 2)
MSR_PLAN(648, 64, 8, 3, 3, 3, 3)
MSR_PLAN(675, 64, 5, 5, 3, 3, 3)
MSR_PLAN(720, 64, 8, 2, 5, 3, 3)
MSR_PLAN(729, 64, 3, 3, 3, 3, 3, 3)
MSR_PLAN(750, 64, 2, 5, 5, 5, 3)
MSR_PLAN(768, 32, 8, 8, 4, 3)
MSR_PLAN(800, 64, 8, 4, 5, 5)
MSR_PLAN(810, 64, 2, 5, 3, 3, 3, 3)
MSR_PLAN(864, 64, 8, 4, 3, 3, 3)
MSR_PLAN(900, 64, 4, 5, 5, 3, 3)
MSR_PLAN(960, 64, 8, 8, 5, 3)
MSR_PLAN(972, 64, 4, 3, 3, 3, 3, 3)
MSR_PLAN(1000, 64, 8, 5, 5, 5)
MSR_PLAN(1024, 64, 8, 8, 4, 4)
MSR_PLAN(1080, 64, 8, 5, 3, 3, 3)
MSR_PLAN(1125, 64, 5, 5, 5, 3, 3)
MSR_PLAN(1152, 64, 8, 8, 2, 3, 3)
MSR_PLAN(1200, 64, 8, 2, 5, 5, 3)
MSR_PLAN(1215, 64, 5, 3, 3, 3, 3, 3)
MSR_PLAN(1250, 64, 2, 5, 5, 5, 5)
MSR_PLAN(1280, 64, 8, 8, 4, 5)
MSR_PLAN(1296, 64, 8, 2, 3, 3, 3, 3)
MSR_PLAN(1350, 64, 2, 5, 5, 3, 3, 3)
MSR_PLAN(1440, 64, 8, 4, 5, 3, 3)
MSR_PLAN(1458, 64, 2, 3, 3, 3, 3, 3, 3)
MSR_PLAN(1500, 64, 4, 5, 5, 5, 3)
MSR_PLAN(1536, 64, 8, 8, 8, 3)
MSR_PLAN(1600, 64, 8, 8, 5, 5)
MSR_PLAN(1620, 64, 4, 5, 3, 3, 3, 3)
MSR_PLAN(1728, 64, 8, 8, 3, 3, 3)
MSR_PLAN(1800, 64, 8, 5, 5, 3, 3)
MSR_PLAN(1875, 64, 5, 5, 5, 5, 3)
MSR_PLAN(1920, 64, 8, 8, 2, 5, 3)
MSR_PLAN(1944, 64, 8, 3, 3, 3, 3, 3)
MSR_PLAN(2000, 64, 8, 2, 5, 5, 5)
MSR_PLAN(2025, 64, 5, 5, 3, 3, 3, 3)
MSR_PLAN(2048, 64, 8, 8, 8, 4)
#undef MSR_PLAN

// Lengths of the fused Toeplitz apply passes (exact plans; the passes are
// instantiated per length, coefficient count and precision).
#define MSR_FFT_LENGTHS(X) \
    X(16) X(24) X(32) X(40) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(320) X(384) \
    X(512) X(640) X(768) X(1024)

// Lengths of the line kernels (every 2,3,5-smooth length from 2 to 2048), in
// 6 parts instantiated by separate translation units (fftlines_p*.hip).
#define MSR_FFT_LINE_LENGTHS_P0(X) X(2) X(9) X(20) X(36) X(60) X(90) X(128) X(180) X(243) X(320) \
    X(405) X(512) X(648) X(800) X(1000) X(1215) X(1458) X(1800) X(2048)
#define MSR_FFT_LINE_LENGTHS_P1(X) X(3) X(10) X(24) X(40) X(64) X(96) X(135) X(192) X(250) \
    X(324) X(432) X(540) X(675) X(810) X(1024) X(1250) X(1500) X(1875)
#define MSR_FFT_LINE_LENGTHS_P2(X) X(4) X(12) X(25) X(45) X(72) X(100) X(144) X(200) X(256) \
    X(360) X(450) X(576) X(720) X(864) X(1080) X(1280) X(1536) X(1920)
#define MSR_FFT_LINE_LENGTHS_P3(X) X(5) X(15) X(27) X(48) X(75) X(108) X(150) X(216) X(270) \
    X(375) X(480) X(600) X(729) X(900) X(1125) X(1296) X(1600) X(1944)
#define MSR_FFT_LINE_LENGTHS_P4(X) X(6) X(16) X(30) X(50) X(80) X(120) X(160) X(225) X(288) \
    X(384) X(486) X(625) X(750) X(960) X(1152) X(1350) X(1620) X(2000)
#define MSR_FFT_LINE_LENGTHS_P5(X) X(8) X(18) X(32) X(54) X(81) X(125) X(162) X(240) X(300) \
    X(400) X(500) X(640) X(768) X(972) X(1200) X(1440) X(1728) X(2025)
#define MSR_FFT_LINE_PARTS 6

#define MSR_FFT_LINE_LENGTHS(X)                                                        \
    MSR_FFT_LINE_LENGTHS_P0(X) MSR_FFT_LINE_LENGTHS_P1(X) MSR_FFT_LINE_LENGTHS_P2(X)         \
    MSR_FFT_LINE_LENGTHS_P3(X) MSR_FFT_LINE_LENGTHS_P4(X) MSR_FFT_LINE_LENGTHS_P5(X)

// Invoke F.template operator()<N>() for a runtime length n of the fused set;
// returns false if n is not one.
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

// The same over every line length (host-side tables: twiddles).
template <class F>
inline bool dispatch_line_len(int n, F&& f) {
#define MSR_CASE(NN)                     \
    case NN:                             \
        f.template operator()<NN>();     \
        return true;
    switch (n) {
        MSR_FFT_LINE_LENGTHS(MSR_CASE)
        default: return false;
    }
#undef MSR_CASE
}

// A length of the fused apply passes.
inline bool fused_len_supported(int64_t n) {
#define MSR_CASE(NN) \
    case NN: return true;
    switch (n) {
        MSR_FFT_LENGTHS(MSR_CASE)
        default: return false;
    }
#undef MSR_CASE
}

// A length of the line kernels (backprojection, forward, kernel build,
// Cartesian operator, the general Toeplitz apply).
inline bool fft_len_supported(int64_t n) {
#define MSR_CASE(NN) \
    case NN: return true;
    switch (n) {
        MSR_FFT_LINE_LENGTHS(MSR_CASE)
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
// Element i of an N-point line lives at lds_pad<N, CT>(i).  The LDS serves a
// wavefront's 16-byte accesses 16 lanes at a time over 16 slots of 16 bytes (a
// probe on MI355X: 16 lanes on 16 distinct slots take the minimum time, two
// lanes per slot twice it), and the transforms access a line in three
// patterns: stride R (the first stage's stores), the blocks of the middle
// stages' stores, and unit stride (every stage's loads, the last stage's
// stores, the kernel multiply and the passes' row copies).
// * ComplexF64 lines of plans with at least 16 threads per line, each thread
//   holding at most 8 values (128, 256, 512 points): element i at
//   i ^ ((i >> 3) & 15), a permutation inside each aligned block of 16, so
//   every one of those patterns of a radix-8 plan meets 16 distinct slots (one
//   complex of padding every 8 elements left the unit-stride accesses at two
//   lanes per slot: measured 38 against 71 probe cycles per wavefront access).
//   Modelled over every plan's stages it would also suit the other plans with
//   16 or more threads per line, but where a thread holds more values the
//   swizzled addresses raise register use until kernels spill (the 320- and
//   384-point z-passes of the 3D apply), so those keep the padding.
// * other ComplexF64 lines: one complex of padding after every 8 elements;
//   ComplexF32: after every 32 (at 256^2 / 20 coils / K = 5, a period of 8 took
//   the ComplexF32 apply to 0.217 ms, 16 to 0.193, 32 to 0.190).
template <class CT>
constexpr int pad_shift() { return sizeof(CT) == 16 ? 3 : 5; }
template <int N, class CT>
constexpr bool lds_swizzled() {
    return sizeof(CT) == 16 && FFTPlan<N>::T >= 16 && N / FFTPlan<N>::T <= 8;
}
// (SWZ: a kernel may keep the padding where the swizzle's address arithmetic
// would make it spill; every access to a line must use the same choice)
template <int N, class CT, bool SWZ = lds_swizzled<N, CT>()>
__host__ __device__ constexpr int lds_pad(int i) {
    if constexpr (SWZ) return i ^ ((i >> 3) & 15);
    else return i + (i >> pad_shift<CT>());
}
template <int N, class CT, bool SWZ = lds_swizzled<N, CT>()>
constexpr int lds_line() {
    if constexpr (SWZ) return (N + 15) & ~15;
    else return N + (N >> pad_shift<CT>());
}

// ----------------------------------------------------------- stages
// The T threads of a line always lie in one wavefront (every plan's T divides
// 64), so the stages of a line are ordered by wavefront-level fences instead
// of workgroup barriers: the LDS operations of a wavefront complete in issue
// order, and the fences keep the compiler from moving LDS accesses across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// True when every stage of the plan gives each thread the same number of
// butterflies (N / R divisible by T): then a thread holds N / T values and the
// first stage's inputs / last stage's outputs can stay in registers.
template <int N>
constexpr bool fft_plan_exact() {
    for (int s = 0; s < FFTPlan<N>::NST; ++s)
        if ((N / FFTPlan<N>::R[s]) % FFTPlan<N>::T != 0) return false;
    return true;
}

// One stage.  MODE bit 0: the inputs come from reg (reg[p R + r] = x[j + r N/R]
// for j = t + p T) instead of LDS; bit 1: the outputs go to reg (for the last
// stage reg[p R + r] = X[j + r N/R]) instead of LDS.
template <int N, int R, int NS, int T, int DIR, int MODE, class CT, bool SWZ>
__device__ __forceinline__ void fft_stage(CT* buf, const CT* tw, int t, bool active, CT* reg) {
    constexpr int NB = N / R;
    constexpr int PER = (NB + T - 1) / T;
    constexpr bool EXACT = (NB % T) == 0;
    constexpr bool IN_REG = (MODE & 1) != 0, OUT_REG = (MODE & 2) != 0;
    static_assert(EXACT || MODE == 0, "register stages need an exact plan");
    CT v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * T;
        if (active && (EXACT || j < NB)) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (IN_REG) v[p][r] = reg[p * R + r];
                else v[p][r] = buf[lds_pad<N, CT, SWZ>(j + r * NB)];
            }
        }
    }
    // every lane's reads of buf (this stage's, or the caller's before a
    // register-input stage) come before any lane's writes
    wave_lds_sync();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * T;
        if (active && (EXACT || j < NB)) {
            const int k = j % NS;
            if constexpr (NS > 1) {
                // stage-ordered table: W_{NS R}^{k r} at OFF + (r - 1) NS + k
                constexpr int OFF = NS - FFTPlan<N>::R[0];
                if constexpr (sizeof(CT) == 16 && R == 8) {
                    // ComplexF64: two table twiddles, the others by products
                    // (rounding <= 3 ulp), fewer LDS reads and registers
                    CT w1 = tw[OFF + k], w4 = tw[OFF + 3 * NS + k];
                    if constexpr (DIR > 0) {
                        w1.y = -w1.y;
                        w4.y = -w4.y;
                    }
                    const CT w2 = cmul(w1, w1), w3 = cmul(w2, w1);
                    v[p][1] = cmul(v[p][1], w1);
                    v[p][2] = cmul(v[p][2], w2);
                    v[p][3] = cmul(v[p][3], w3);
                    v[p][4] = cmul(v[p][4], w4);
                    v[p][5] = cmul(v[p][5], cmul(w4, w1));
                    v[p][6] = cmul(v[p][6], cmul(w4, w2));
                    v[p][7] = cmul(v[p][7], cmul(w4, w3));
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        CT w = tw[OFF + (r - 1) * NS + k];
                        if constexpr (DIR > 0) w.y = -w.y;
                        v[p][r] = cmul(v[p][r], w);
                    }
                }
            }
            Butterfly<R, DIR>::run(v[p]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (OUT_REG) reg[p * R + r] = v[p][r];
                else buf[lds_pad<N, CT, SWZ>(base + r * NS)] = v[p][r];
            }
        }
    }
    if constexpr (!OUT_REG) wave_lds_sync();  // visible to the next stage's reads
}

template <int N, int T, int DIR, int S, int NS, int MODE, class CT, bool SWZ>
__device__ __forceinline__ void fft_stages(CT* buf, const CT* tw, int t, bool active, CT* reg) {
    constexpr int NST = FFTPlan<N>::NST;
    if constexpr (S < NST) {
        constexpr int R = FFTPlan<N>::R[S];
        constexpr int M = (S == 0 ? (MODE & 1) : 0) | (S == NST - 1 ? (MODE & 2) : 0);
        fft_stage<N, R, NS, T, DIR, M, CT, SWZ>(buf, tw, t, active, reg);
        fft_stages<N, T, DIR, S + 1, NS * R, MODE, CT, SWZ>(buf, tw, t, active, reg);
    }
}

// In-place N-point FFT of a line (element i at buf[lds_pad<N, CT>(i)]) by the
// T = FFTPlan<N>::T threads with local index t.  Requires a barrier before the
// call (data in LDS); ends with a barrier.  All threads of the workgroup must
// call it; threads with active = false only take part in the final barrier
// (buf is not touched), for workgroups with more threads than line work.
template <int N, int DIR, class CT, bool SWZ = lds_swizzled<N, CT>()>
__device__ __forceinline__ void fft_lds(CT* buf, const CT* tw, int t, bool active = true) {
    fft_stages<N, FFTPlan<N>::T, DIR, 0, 1, 0, CT, SWZ>(buf, tw, t, active, nullptr);
    __syncthreads();
}

// The same transform with no workgroup barrier, for an exact plan, the line's
// buf used only by its own wavefront.  IN_REG: the input is in reg in the
// first stage's layout, reg[p R0 + r] = x[t + p T + r N/R0] (otherwise in buf,
// ordered by the caller); OUT_REG: the output is left in reg in the last
// stage's layout, reg[p RL + r] = X[t + p T + r N/RL] (otherwise in buf in
// natural order, visible to this wavefront; other wavefronts need a barrier).
// reg holds N / T values.
template <int N, int DIR, bool IN_REG, bool OUT_REG, class CT, bool SWZ = lds_swizzled<N, CT>()>
__device__ __forceinline__ void fft_wave(CT* buf, const CT* tw, int t, bool active, CT* reg) {
    static_assert(fft_plan_exact<N>(), "fft_wave needs an exact plan");
    fft_stages<N, FFTPlan<N>::T, DIR, 0, 1, (IN_REG ? 1 : 0) | (OUT_REG ? 2 : 0), CT, SWZ>(buf, tw, t, active, reg);
}

// Copy an N-entry twiddle table from global memory into LDS (no barrier).
template <int N, class CT>
__device__ __forceinline__ void load_twiddles(CT* s_tw, const CT* g_tw) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) s_tw[i] = g_tw[i];
}

}  // namespace msr
