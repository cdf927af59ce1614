// ionp_common.h -- shared definitions for the gfx950 ufunc/reduce backend:
// dtype traits, NumPy-on-x86 scalar semantics (device + host), kernel
// operand descriptors and a 32-bit fast divider.
//
// Scalar semantics follow the reference's scalar kernels and what its NumPy
// oracle returns on x86-64:
//   float_same  ionp-core/src/ufunc.rs:2460-2579  (IEEE + - * /, no FMA)
//   int_same    ionp-core/src/ufunc.rs:2305       (wrapping integers)
//   float_max   ionp-core/src/ufunc.rs:2402-2429  (NaN-propagating max/min)
//   cmp_same    ionp-core/src/ufunc.rs:3479       (IEEE compare, NaN -> false)
//   casts       ionp-core/src/buffer.rs:81-210
// NaN results: the first NaN operand, quieted; an invalid operation yields
// the x86 default NaN (sign bit set).  max/min return a NaN operand as is and
// the second operand on ties (-0 vs +0).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <type_traits>

#include "../../include/ionp_hip.h"

namespace ionp {

constexpr int kMaxKDims = 8;  // rank after coalescing handled by kernels

enum : int32_t {
    OP_WHERE = 100,  // elementwise kernel op codes beyond ionp_op
    OP_CAST = 101,
};

__host__ __device__ constexpr int itemsize(int dt) {
    return dt == IONP_BOOL || dt == IONP_INT8 || dt == IONP_UINT8                    ? 1
           : dt == IONP_INT16 || dt == IONP_UINT16 || dt == IONP_FLOAT16              ? 2
           : dt == IONP_INT32 || dt == IONP_UINT32 || dt == IONP_FLOAT32              ? 4
                                                                                        : 8;
}
// float32 / float64: the dtypes the native float kernels compute in
__host__ __device__ constexpr bool is_float_dt(int dt) { return dt == IONP_FLOAT32 || dt == IONP_FLOAT64; }
// any floating dtype, float16 included (dtype rules)
__host__ __device__ constexpr bool is_real_float(int dt) { return is_float_dt(dt) || dt == IONP_FLOAT16; }
__host__ __device__ constexpr bool is_unsigned_dt(int dt) {
    return dt == IONP_UINT8 || dt == IONP_UINT16 || dt == IONP_UINT32 || dt == IONP_UINT64;
}
__host__ __device__ constexpr bool is_signed_dt(int dt) {
    return dt == IONP_INT8 || dt == IONP_INT16 || dt == IONP_INT32 || dt == IONP_INT64;
}

template <int DT> struct dt_type;
template <> struct dt_type<IONP_BOOL> { using T = uint8_t; };
template <> struct dt_type<IONP_INT8> { using T = int8_t; };
template <> struct dt_type<IONP_UINT8> { using T = uint8_t; };
template <> struct dt_type<IONP_INT16> { using T = int16_t; };
template <> struct dt_type<IONP_UINT16> { using T = uint16_t; };
template <> struct dt_type<IONP_INT32> { using T = int32_t; };
template <> struct dt_type<IONP_UINT32> { using T = uint32_t; };
template <> struct dt_type<IONP_INT64> { using T = int64_t; };
template <> struct dt_type<IONP_UINT64> { using T = uint64_t; };
template <> struct dt_type<IONP_FLOAT32> { using T = float; };
template <> struct dt_type<IONP_FLOAT64> { using T = double; };

// ---------------------------------------------------------------------------
// bit helpers and NaN rules
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
__host__ __device__ inline uint32_t bits(float x) { return __builtin_bit_cast(uint32_t, x); }
__host__ __device__ inline double f64_of(uint64_t u) { return __builtin_bit_cast(double, u); }
__host__ __device__ inline float f32_of(uint32_t u) { return __builtin_bit_cast(float, u); }

__host__ __device__ inline double quiet(double x) { return f64_of(bits(x) | 0x0008000000000000ull); }
__host__ __device__ inline float quiet(float x) { return f32_of(bits(x) | 0x00400000u); }
__host__ __device__ inline double default_nan(double) { return f64_of(0xfff8000000000000ull); }
__host__ __device__ inline float default_nan(float) { return f32_of(0xffc00000u); }
__host__ __device__ inline double canonical_nan(double) { return f64_of(0x7ff8000000000000ull); }
__host__ __device__ inline float canonical_nan(float) { return f32_of(0x7fc00000u); }
template <typename T> __host__ __device__ inline bool isnan_(T x) { return x != x; }

// ---------------------------------------------------------------------------
// float16 (NumPy's npy_half): conversions as NumPy 2.2.6 performs them in
// software (npy_half_to_float / npy_float_to_half / npy_double_to_half):
// round to nearest even, NaN payloads kept (top significand bits, never
// quieted, an all-zero payload becomes 1), overflow (a finite value
// rounding to inf) and underflow (a result below the smallest normal half
// that is not exact -- tininess before rounding) reported in *fl as the
// IEEE flag bits 2 / 4.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t f16_to_f32_bits(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = h & 0x7c00u, m = h & 0x03ffu;
    if (e == 0x7c00u) return sign | 0x7f800000u | (m << 13);  // inf / NaN, payload kept
    if (e != 0) return sign | (((e >> 10) + 112u) << 23) | (m << 13);
    if (m == 0) return sign;
    // subnormal half: m * 2^-24, normalised
    int sh = 0;
    uint32_t mm = m;
    while (!(mm & 0x0400u)) {
        mm <<= 1;
        sh++;
    }
    return sign | ((uint32_t)(113 - sh) << 23) | ((mm & 0x03ffu) << 13);
}
__host__ __device__ inline uint64_t f16_to_f64_bits(uint16_t h) {
    const uint64_t sign = (uint64_t)(h & 0x8000u) << 48;
    const uint32_t e = h & 0x7c00u, m = h & 0x03ffu;
    if (e == 0x7c00u) return sign | 0x7ff0000000000000ull | ((uint64_t)m << 42);
    const uint32_t fb = f16_to_f32_bits(h);  // finite: exact through float
    const double d = (double)__builtin_bit_cast(float, fb);
    return __builtin_bit_cast(uint64_t, d) | (m == 0 && e == 0 ? sign : 0);
}
// round the significand `sig` (implicit bit included) right by `drop` bits,
// nearest even; *inexact set when bits were lost
__host__ __device__ inline uint64_t rne_shift(uint64_t sig, int drop, bool *inexact) {
    if (drop <= 0) {
        *inexact = false;
        return sig << (-drop);
    }
    if (drop >= 64) {
        *inexact = sig != 0;
        return 0;
    }
    const uint64_t q = sig >> drop, rem = sig & ((1ull << drop) - 1), half = 1ull << (drop - 1);
    *inexact = rem != 0;
    return q + ((rem > half || (rem == half && (q & 1))) ? 1 : 0);
}
// value = sig * 2^(e2) (sig with its implicit bit, e2 the exponent of the
// significand's last bit) to half bits
__host__ __device__ inline uint16_t f16_round(uint32_t sign, uint64_t sig, int e2, int msb, uint32_t *fl) {
    // msb: index of sig's top bit; the value's binary exponent is e2 + msb
    const int ex = e2 + msb;
    if (ex >= 16) {  // >= 2^16: inf
        *fl |= 2;
        return (uint16_t)(sign | 0x7c00u);
    }
    bool inexact;
    if (ex < -14) {  // half subnormal range: units of 2^-24
        const uint64_t q = rne_shift(sig, -24 - e2, &inexact);
        if (inexact) *fl |= 4;
        return (uint16_t)(sign | (uint32_t)q);  // q == 0x400 is the smallest normal
    }
    const uint64_t q = rne_shift(sig, msb - 10, &inexact);  // 11 significant bits
    const uint32_t h = (uint32_t)((ex + 15) << 10) + (uint32_t)(q - 0x400u);  // a carry moves into the exponent
    if (h >= 0x7c00u) {
        *fl |= 2;
        return (uint16_t)(sign | 0x7c00u);
    }
    return (uint16_t)(sign | h);
}
__host__ __device__ inline uint16_t f32_to_f16_bits(uint32_t f, uint32_t *fl) {
    const uint32_t sign = (f >> 16) & 0x8000u, a = f & 0x7fffffffu;
    if (a >= 0x7f800000u) {
        if (a == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
        uint32_t r = 0x7c00u + ((a & 0x007fffffu) >> 13);
        if (r == 0x7c00u) r++;
        return (uint16_t)(sign | r);
    }
    if (a == 0) return (uint16_t)sign;
    const int e = (int)(a >> 23);
    uint64_t sig = a & 0x007fffffu;
    int e2;
    if (e == 0) {
        e2 = -149;  // float subnormal
    } else {
        sig |= 0x00800000u;
        e2 = e - 150;
    }
    const int msb = 63 - __builtin_clzll(sig);
    if (e2 + msb < -25) {  // far below half's range: zero, underflowed
        *fl |= 4;
        return (uint16_t)sign;
    }
    return f16_round(sign, sig, e2, msb, fl);
}
__host__ __device__ inline uint16_t f64_to_f16_bits(uint64_t d, uint32_t *fl) {
    const uint32_t sign = (uint32_t)(d >> 48) & 0x8000u;
    const uint64_t a = d & 0x7fffffffffffffffull;
    if (a >= 0x7ff0000000000000ull) {
        if (a == 0x7ff0000000000000ull) return (uint16_t)(sign | 0x7c00u);
        uint32_t r = 0x7c00u + (uint32_t)((a & 0x000fffffffffffffull) >> 42);
        if (r == 0x7c00u) r++;
        return (uint16_t)(sign | r);
    }
    if (a == 0) return (uint16_t)sign;
    const int e = (int)(a >> 52);
    uint64_t sig = a & 0x000fffffffffffffull;
    int e2;
    if (e == 0) {
        e2 = -1074;
    } else {
        sig |= 0x0010000000000000ull;
        e2 = e - 1075;
    }
    const int msb = 63 - __builtin_clzll(sig);
    if (e2 + msb < -25) {
        *fl |= 4;
        return (uint16_t)sign;
    }
    return f16_round(sign, sig, e2, msb, fl);
}
__host__ __device__ inline float h2f(uint16_t h) { return __builtin_bit_cast(float, f16_to_f32_bits(h)); }
__host__ __device__ inline double h2d(uint16_t h) { return __builtin_bit_cast(double, f16_to_f64_bits(h)); }
__host__ __device__ inline uint16_t f2h(float x, uint32_t *fl) { return f32_to_f16_bits(__builtin_bit_cast(uint32_t, x), fl); }
__host__ __device__ inline uint16_t d2h(double x, uint32_t *fl) { return f64_to_f16_bits(__builtin_bit_cast(uint64_t, x), fl); }
__host__ __device__ inline bool h_isnan(uint16_t h) { return (h & 0x7c00u) == 0x7c00u && (h & 0x03ffu); }

// Slow path of an arithmetic op whose hardware result was NaN.
template <typename T> __host__ __device__ __noinline__ T nan_result(T a, T b) {
    if (isnan_(a)) return quiet(a);
    if (isnan_(b)) return quiet(b);
    return default_nan(a);
}

template <typename T> __host__ __device__ inline T fadd(T a, T b) {
    T r = a + b;
    return r == r ? r : nan_result(a, b);
}
template <typename T> __host__ __device__ inline T fsub(T a, T b) {
    T r = a - b;
    return r == r ? r : nan_result(a, b);
}
template <typename T> __host__ __device__ inline T fmul(T a, T b) {
    T r = a * b;
    return r == r ? r : nan_result(a, b);
}
template <typename T> __host__ __device__ inline T fdiv(T a, T b) {
    T r = a / b;
    return r == r ? r : nan_result(a, b);
}
// maximum / minimum: first-operand NaN, then second-operand NaN, ties -> b.
template <typename T> __host__ __device__ inline T fmaxn(T a, T b) {
    return isnan_(a) ? a : (isnan_(b) ? b : (a > b ? a : b));
}
template <typename T> __host__ __device__ inline T fminn(T a, T b) {
    return isnan_(a) ? a : (isnan_(b) ? b : (a < b ? a : b));
}

// ---------------------------------------------------------------------------
// casts (x86 results for NaN payloads and out-of-range float -> int)
// ---------------------------------------------------------------------------
__host__ __device__ inline int32_t cvtt_i32(double x) {
    if (!(x < 2147483648.0 && x > -2147483649.0)) return INT32_MIN;  // NaN too
    return (int32_t)x;
}
__host__ __device__ inline int64_t cvtt_i64(double x) {
    if (!(x < 9223372036854775808.0 && x >= -9223372036854775808.0)) return INT64_MIN;
    return (int64_t)x;
}
__host__ __device__ inline uint64_t cvtt_u64(double x) {
    if (x >= 9223372036854775808.0) return (uint64_t)cvtt_i64(x - 9223372036854775808.0) ^ 0x8000000000000000ull;
    return (uint64_t)cvtt_i64(x);
}
__host__ __device__ inline double widen_f32(float f) {
    if (f != f) {
        uint32_t b = bits(f) | 0x00400000u;
        return f64_of(((uint64_t)(b >> 31) << 63) | 0x7ff0000000000000ull | ((uint64_t)(b & 0x7fffffu) << 29));
    }
    return (double)f;
}
__host__ __device__ inline float narrow_f64(double x) {
    if (x != x) {
        uint64_t b = bits(x) | 0x0008000000000000ull;
        return f32_of((uint32_t)((b >> 63) << 31) | 0x7f800000u | (uint32_t)((b >> 29) & 0x7fffffu));
    }
    return (float)x;
}

// Convert a value of C type S to the C type of dtype D with NumPy's rules.
template <int D, typename S> __host__ __device__ inline typename dt_type<D>::T cvt(S v) {
    using TD = typename dt_type<D>::T;
    if constexpr (D == IONP_BOOL) {
        return (TD)(v != (S)0);
    } else if constexpr (std::is_same<S, float>::value || std::is_same<S, double>::value) {
        double x;
        if constexpr (std::is_same<S, float>::value) {
            if constexpr (D == IONP_FLOAT32) return v;
            x = widen_f32(v);
        } else {
            x = v;
        }
        if constexpr (D == IONP_FLOAT64) return x;
        else if constexpr (D == IONP_FLOAT32) return narrow_f64(x);
        else if constexpr (D == IONP_INT8 || D == IONP_UINT8 || D == IONP_INT16 || D == IONP_UINT16)
            return (TD)(uint32_t)cvtt_i32(x);
        else if constexpr (D == IONP_INT32) return cvtt_i32(x);
        else if constexpr (D == IONP_UINT32) return (TD)(uint64_t)cvtt_i64(x);
        else if constexpr (D == IONP_INT64) return cvtt_i64(x);
        else return cvtt_u64(x);
    } else {
        // integer (or bool) source: C conversions (RNE for int -> float,
        // two's-complement truncation between integers)
        return (TD)v;
    }
}

// ---------------------------------------------------------------------------
// 32-bit fast division by a runtime-invariant divisor (d >= 1, n < 2^31):
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
// ---------------------------------------------------------------------------
struct FastDiv {
    uint32_t d, m, s;
    FastDiv() = default;
    __host__ explicit FastDiv(uint32_t dv) {
        d = dv;
        s = 0;
        while ((1ull << s) < dv) ++s;
        m = (uint32_t)((((1ull << s) - dv) << 32) / dv + 1);
    }
    __host__ __device__ inline uint32_t div(uint32_t n) const {
#if defined(__HIP_DEVICE_COMPILE__)
        uint32_t hi = __umulhi(n, m);
#else
        uint32_t hi = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
        return (uint32_t)(((uint64_t)hi + n) >> s);
    }
};

// ---------------------------------------------------------------------------
// elementwise kernel descriptor: up to 3 inputs + 1 output over a coalesced
// iteration space of rank <= kMaxKDims (dim 0 innermost)
// ---------------------------------------------------------------------------
struct EwArgs {
    int32_t op;        // ionp_op, OP_WHERE or OP_CAST
    int32_t nin;
    int32_t ndim;
    int32_t in_dt[3];
    int32_t out_dt;
    int32_t loop_dt;
    int32_t mixed_su;  // int64-vs-uint64 exact comparison
    int64_t shape[kMaxKDims];
    FastDiv div[kMaxKDims];
    int64_t in_st[3][kMaxKDims];  // elements
    int64_t out_st[kMaxKDims];
    const char *in[3];            // address of element 0
    char *out;
    int64_t total;
    uint32_t *fpe;                // IEEE flag word to OR into, or null (not tracked)
};

}  // namespace ionp
