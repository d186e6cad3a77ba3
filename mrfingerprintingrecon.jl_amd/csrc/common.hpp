// Shared helpers for the MI355X (gfx950) library: status codes, error state,
// complex arithmetic on float2/double2, HIP error checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "msr.h"

namespace msr {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

#define MSR_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            ::msr::fail(e_ == hipErrorOutOfMemory ? MSR_ERR_NOMEM : MSR_ERR_HIP,         \
                        "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                       \
    } while (0)

#define MSR_CHECK_LAUNCH() MSR_HIP(hipGetLastError())

#define MSR_REQUIRE(cond, code, ...)          \
    do {                                      \
        if (!(cond)) ::msr::fail(code, __VA_ARGS__); \
    } while (0)

void set_last_error(const std::string& s);

// ---------------------------------------------------------------- complex
__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// acc += a * b
__host__ __device__ __forceinline__ void cfma(float2& acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
}
// acc += conj(a) * b
__host__ __device__ __forceinline__ void cfmac(float2& acc, float2 a, float2 b) {
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(-a.y, b.x, acc.y);
}

// double-precision complex (same operations)
__host__ __device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
__host__ __device__ __forceinline__ void cfmac(double2& acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
}

// Complex type traits: CT is float2 (ComplexF32) or double2 (ComplexF64).
template <class CT> struct CTraits;
template <> struct CTraits<float2> {
    using R = float;
    __host__ __device__ static float2 make(double a, double b) { return make_float2((float)a, (float)b); }
};
template <> struct CTraits<double2> {
    using R = double;
    __host__ __device__ static double2 make(double a, double b) { return make_double2(a, b); }
};
template <class CT> using real_t = typename CTraits<CT>::R;
template <class CT>
__host__ __device__ __forceinline__ CT cz() { return CTraits<CT>::make(0.0, 0.0); }
// promote / convert a ComplexF32 value
template <class CT>
__host__ __device__ __forceinline__ CT cto(float2 a) { return CTraits<CT>::make(a.x, a.y); }
template <class CT>
__host__ __device__ __forceinline__ CT cto(double2 a) { return CTraits<CT>::make(a.x, a.y); }

// ---------------------------------------------------------------- storage
// Compact storage of a ComplexF64 value in 64 bits (the FFT intermediates of
// the MSR_PREC_F64C apply, whose arithmetic is FP64): one binary exponent
// shared by the two parts and two 28-bit two's-complement mantissas.  With
// m = max(|re|, |im|) in [2^(E-1), 2^E), each part is rounded to a multiple of
// 2^(E-27): error <= 2^-27 m per part, 8x below ComplexF32 storage
// (2^-24 |part|) in the same 8 bytes.  Layout: w0 = [E + 128 : 8 | re >> 4 :
// 24], w1 = [re & 15 : 4 | im : 28]; exponent byte 0 is zero (also for values
// below 2^-128), 255 is not finite or at least 2^126 (read back as NaN).
struct __align__(8) cx64 {
    uint32_t w0, w1;
};

// Integer and bit operations on the FP64 words (the exponent from the larger
// biased exponent field, the power-of-two scales built from bits, rounding
// to the nearest integer by adding 1.5 * 2^52 and keeping the low word), so
// the conversion costs two FP64 FMAs (encode) or two conversions and
// multiplies (decode) besides 32-bit integer operations.
__device__ __forceinline__ cx64 cx64_encode(double re, double im) {
    const uint32_t hr = (uint32_t)(__double_as_longlong(re) >> 32);
    const uint32_t hi = (uint32_t)(__double_as_longlong(im) >> 32);
    const uint32_t er = (hr >> 20) & 0x7FFu, ei = (hi >> 20) & 0x7FFu;
    const uint32_t eb = er > ei ? er : ei;  // biased exponent of max(|re|, |im|)
    // max(|re|, |im|) in [2^(eb-1023), 2^(eb-1022)): stored exponent
    // E = (eb - 1022) + 128, parts scaled by 2^(27 - (eb - 1022)) = 2^(1049 - eb)
    cx64 r;
    if (eb < 895u) {  // below 2^-128 (and zero): stored as zero
        r.w0 = 0u;
        r.w1 = 0u;
        return r;
    }
    if (eb > 1148u) {  // 2^126 and above, infinite or NaN
        r.w0 = 0xFF000000u;
        r.w1 = 0u;
        return r;
    }
    const double s = __longlong_as_double((long long)(2072u - eb) << 52);  // 2^(1049 - eb)
    constexpr double magic = 6755399441055744.0;                          // 1.5 * 2^52
    // |part * s| < 2^27: the low word of part * s + 1.5 2^52 is the rounded
    // part (two's complement); a part that rounds to 2^27 is clamped
    int qr = (int)(uint32_t)__double_as_longlong(fma(re, s, magic));
    int qi = (int)(uint32_t)__double_as_longlong(fma(im, s, magic));
    const int lim = (1 << 27) - 1;
    qr = qr > lim ? lim : (qr < -lim ? -lim : qr);
    qi = qi > lim ? lim : (qi < -lim ? -lim : qi);
    r.w0 = ((eb - 894u) << 24) | (((uint32_t)qr >> 4) & 0x00FFFFFFu);
    r.w1 = ((uint32_t)qr << 28) | ((uint32_t)qi & 0x0FFFFFFFu);
    return r;
}

__device__ __forceinline__ double2 cx64_decode(cx64 v) {
    const uint32_t E = v.w0 >> 24;
    const int qr = ((int)(v.w0 << 8) >> 4) | (int)(v.w1 >> 28);
    const int qi = (int)(v.w1 << 4) >> 4;
    // scale 2^(E - 128 - 27): biased exponent field E + 868 (E = 0 stores
    // zero parts, so the tiny scale gives 0)
    const double s = __longlong_as_double((long long)(E + 868u) << 52);
    double2 r = make_double2((double)qr * s, (double)qi * s);
    if (E == 255u) r = make_double2(__builtin_nan(""), __builtin_nan(""));
    return r;
}

// Load / store of an intermediate of storage type ST in arithmetic type CT:
// ST == CT (float2 or double2 in memory), or ST == cx64 with CT == double2.
template <class CT, class ST>
__device__ __forceinline__ CT st_load(const ST* p) {
    if constexpr (std::is_same<ST, cx64>::value) return cx64_decode(*p);
    else return *p;
}
template <class CT, class ST>
__device__ __forceinline__ void st_store(ST* p, CT v) {
    if constexpr (std::is_same<ST, cx64>::value) *p = cx64_encode(v.x, v.y);
    else *p = v;
}
template <class CT, class ST>
__device__ __forceinline__ CT st_conv(ST v) {
    if constexpr (std::is_same<ST, cx64>::value) return cx64_decode(v);
    else return v;
}

// Opaque copy of a register value: work on the copy cannot be hoisted out of
// the enclosing loop (keeps a register-resident Float32 table from being
// widened to Float64 once for all iterations, which doubles its registers).
__device__ __forceinline__ float opaque(float v) {
    asm volatile("" : "+v"(v));
    return v;
}
__device__ __forceinline__ float2 opaque(float2 v) { return make_float2(opaque(v.x), opaque(v.y)); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace msr
