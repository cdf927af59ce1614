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
    return dt == IONP_BOOL || dt == IONP_INT8 || dt == IONP_UINT8     ? 1
           : dt == IONP_INT16 || dt == IONP_UINT16                    ? 2
           : dt == IONP_INT32 || dt == IONP_UINT32 || dt == IONP_FLOAT32 ? 4
                                                                         : 8;
}
__host__ __device__ constexpr bool is_float_dt(int dt) { return dt == IONP_FLOAT32 || dt == IONP_FLOAT64; }
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
