// device_ops.h -- device-side building blocks shared by the elementwise,
// reduction and batch kernels: NumPy-semantics scalar ops on a loop dtype,
// dynamic-dtype loads/stores, the per-element body of the generic
// elementwise kernel, reduction index helpers and NumPy's pairwise tree
// evaluated by one workgroup.
#pragma once

#include "elementwise.h"
#include "reduce.h"

namespace ionp {

// ---------------------------------------------------------------------------
// scalar ops on the loop dtype
// ---------------------------------------------------------------------------
template <typename T> struct Arith {
    static constexpr bool is_f = std::is_floating_point<T>::value;
    using U = typename std::make_unsigned<typename std::conditional<is_f, int, T>::type>::type;
    __device__ static inline T add(T a, T b) {
        if constexpr (is_f) return fadd(a, b);
        else return (T)((U)a + (U)b);
    }
    __device__ static inline T sub(T a, T b) {
        if constexpr (is_f) return fsub(a, b);
        else return (T)((U)a - (U)b);
    }
    __device__ static inline T mul(T a, T b) {
        if constexpr (is_f) return fmul(a, b);
        else return (T)((U)a * (U)b);
    }
    __device__ static inline T div(T a, T b) {
        if constexpr (is_f) return fdiv(a, b);
        else return (T)0;  // integer true_divide resolves to a float loop
    }
    __device__ static inline T max(T a, T b) {
        if constexpr (is_f) return fmaxn(a, b);
        else return a > b ? a : b;
    }
    __device__ static inline T min(T a, T b) {
        if constexpr (is_f) return fminn(a, b);
        else return a < b ? a : b;
    }
};

// bool loop (ufunc.rs bool kernels): add/maximum = or, multiply/minimum = and
template <int LT, typename T = typename dt_type<LT>::T> __device__ inline T arith_op(int op, T a, T b) {
    if constexpr (LT == IONP_BOOL) {
        const uint8_t x = a != 0, y = b != 0;
        return (op == IONP_ADD || op == IONP_MAXIMUM) ? (T)(x | y) : (T)(x & y);
    } else {
        switch (op) {
            case IONP_ADD: return Arith<T>::add(a, b);
            case IONP_SUBTRACT: return Arith<T>::sub(a, b);
            case IONP_MULTIPLY: return Arith<T>::mul(a, b);
            case IONP_TRUE_DIVIDE: return Arith<T>::div(a, b);
            case IONP_MAXIMUM: return Arith<T>::max(a, b);
            default: return Arith<T>::min(a, b);
        }
    }
}

template <typename T> __device__ inline bool cmp_op(int op, T a, T b) {
    switch (op) {
        case IONP_LESS: return a < b;
        case IONP_LESS_EQUAL: return a <= b;
        case IONP_GREATER: return a > b;
        case IONP_GREATER_EQUAL: return a >= b;
        case IONP_EQUAL: return a == b;
        default: return a != b;
    }
}

// exact int64 vs uint64 comparison (NumPy 2's mixed-sign comparison loop)
__device__ inline bool cmp_su(int op, int64_t a, uint64_t b) {
    const bool lt = a < 0 || (uint64_t)a < b, eq = a >= 0 && (uint64_t)a == b;
    switch (op) {
        case IONP_LESS: return lt;
        case IONP_LESS_EQUAL: return lt || eq;
        case IONP_GREATER: return !lt && !eq;
        case IONP_GREATER_EQUAL: return !lt;
        case IONP_EQUAL: return eq;
        default: return !eq;
    }
}
__device__ inline int swap_cmp(int op) {
    switch (op) {
        case IONP_LESS: return IONP_GREATER;
        case IONP_LESS_EQUAL: return IONP_GREATER_EQUAL;
        case IONP_GREATER: return IONP_LESS;
        case IONP_GREATER_EQUAL: return IONP_LESS_EQUAL;
        default: return op;
    }
}

// dynamic-dtype load converted to the loop dtype
template <int LT> __device__ inline typename dt_type<LT>::T load_as(int dt, const char *p) {
    switch (dt) {
        case IONP_BOOL: return cvt<LT>((uint8_t)(*(const uint8_t *)p != 0));
        case IONP_INT8: return cvt<LT>(*(const int8_t *)p);
        case IONP_UINT8: return cvt<LT>(*(const uint8_t *)p);
        case IONP_INT16: return cvt<LT>(*(const int16_t *)p);
        case IONP_UINT16: return cvt<LT>(*(const uint16_t *)p);
        case IONP_INT32: return cvt<LT>(*(const int32_t *)p);
        case IONP_UINT32: return cvt<LT>(*(const uint32_t *)p);
        case IONP_INT64: return cvt<LT>(*(const int64_t *)p);
        case IONP_UINT64: return cvt<LT>(*(const uint64_t *)p);
        case IONP_FLOAT32: return cvt<LT>(*(const float *)p);
        default: return cvt<LT>(*(const double *)p);
    }
}

// dynamic-dtype store of a loop-dtype value
template <typename T> __device__ inline void store_as(int dt, char *p, T v) {
    switch (dt) {
        case IONP_BOOL: *(uint8_t *)p = cvt<IONP_BOOL>(v); break;
        case IONP_INT8: *(int8_t *)p = cvt<IONP_INT8>(v); break;
        case IONP_UINT8: *(uint8_t *)p = cvt<IONP_UINT8>(v); break;
        case IONP_INT16: *(int16_t *)p = cvt<IONP_INT16>(v); break;
        case IONP_UINT16: *(uint16_t *)p = cvt<IONP_UINT16>(v); break;
        case IONP_INT32: *(int32_t *)p = cvt<IONP_INT32>(v); break;
        case IONP_UINT32: *(uint32_t *)p = cvt<IONP_UINT32>(v); break;
        case IONP_INT64: *(int64_t *)p = cvt<IONP_INT64>(v); break;
        case IONP_UINT64: *(uint64_t *)p = cvt<IONP_UINT64>(v); break;
        case IONP_FLOAT32: *(float *)p = cvt<IONP_FLOAT32>(v); break;
        default: *(double *)p = cvt<IONP_FLOAT64>(v); break;
    }
}

// the value of an item whose raw bytes are the low bytes of b (little
// endian), converted to the loop dtype as load_as converts
template <int LT> __device__ inline typename dt_type<LT>::T from_bits(int dt, uint64_t b) {
    switch (dt) {
        case IONP_BOOL: return cvt<LT>((uint8_t)((uint8_t)b != 0));
        case IONP_INT8: return cvt<LT>((int8_t)(uint8_t)b);
        case IONP_UINT8: return cvt<LT>((uint8_t)b);
        case IONP_INT16: return cvt<LT>((int16_t)(uint16_t)b);
        case IONP_UINT16: return cvt<LT>((uint16_t)b);
        case IONP_INT32: return cvt<LT>((int32_t)(uint32_t)b);
        case IONP_UINT32: return cvt<LT>((uint32_t)b);
        case IONP_INT64: return cvt<LT>((int64_t)b);
        case IONP_UINT64: return cvt<LT>((uint64_t)b);
        case IONP_FLOAT32: return cvt<LT>(__builtin_bit_cast(float, (uint32_t)b));
        default: return cvt<LT>(__builtin_bit_cast(double, b));
    }
}

// Operand sources of one element: operands read from their addresses, or
// held as raw bit patterns in registers (loaded ahead by the caller).
struct PtrSrc {
    const char *p[3];
    template <int L> __device__ typename dt_type<L>::T val(int k, int dt) const { return load_as<L>(dt, p[k]); }
    __device__ int64_t i64(int k) const { return *(const int64_t *)p[k]; }
    __device__ uint64_t u64(int k) const { return *(const uint64_t *)p[k]; }
    // a float32 / float64 operand as double
    __device__ double fx(int k, int dt) const { return dt == IONP_FLOAT32 ? (double)*(const float *)p[k] : *(const double *)p[k]; }
};
struct BitSrc {
    uint64_t b[3];
    template <int L> __device__ typename dt_type<L>::T val(int k, int dt) const { return from_bits<L>(dt, b[k]); }
    __device__ int64_t i64(int k) const { return (int64_t)b[k]; }
    __device__ uint64_t u64(int k) const { return b[k]; }
    __device__ double fx(int k, int dt) const {
        return dt == IONP_FLOAT32 ? (double)__builtin_bit_cast(float, (uint32_t)b[k]) : __builtin_bit_cast(double, b[k]);
    }
};

// One output element of the generic kernel from its operand source.
template <int LT, class S> __device__ inline void ew_element_src(const EwArgs &A, const S &s, char *po) {
    using T = typename dt_type<LT>::T;
    if (A.op == OP_CAST) {
        store_as<T>(A.out_dt, po, s.template val<LT>(0, A.in_dt[0]));
    } else if (A.op == OP_WHERE) {
        const bool c = s.template val<IONP_BOOL>(0, A.in_dt[0]) != 0;
        store_as<T>(A.out_dt, po, c ? s.template val<LT>(1, A.in_dt[1]) : s.template val<LT>(2, A.in_dt[2]));
    } else if (A.op >= IONP_LESS) {
        bool r;
        if (A.mixed_su) {
            r = A.in_dt[0] == IONP_INT64 ? cmp_su(A.op, s.i64(0), s.u64(1)) : cmp_su(swap_cmp(A.op), s.i64(1), s.u64(0));
        } else {
            r = cmp_op<T>(A.op, s.template val<LT>(0, A.in_dt[0]), s.template val<LT>(1, A.in_dt[1]));
        }
        *(uint8_t *)po = (uint8_t)r;
    } else {
        store_as<T>(A.out_dt, po,
                    arith_op<LT>(A.op, s.template val<LT>(0, A.in_dt[0]), s.template val<LT>(1, A.in_dt[1])));
    }
}

// One output element of the generic kernel, operand addresses resolved.
template <int LT>
__device__ inline void ew_element(const EwArgs &A, const char *p0, const char *p1, const char *p2, char *po) {
    ew_element_src<LT>(A, PtrSrc{{p0, p1, p2}}, po);
}

// ---------------------------------------------------------------------------
// IEEE exception flags, as NumPy reads them from the x86 status register
// after a loop (np.errstate categories: divide 1, over 2, under 4, invalid
// 8).  Derived per element from the operands and the result:
//   invalid  a NaN result from non-NaN operands (inf - inf, 0 * inf, 0 / 0,
//            inf / inf)
//   divide   finite nonzero / 0
//   over     an infinite result from finite operands
//   under    tiny after rounding (x86 rule) and inexact; sums are exact
//            when tiny, so only products and quotients underflow
// Signalling NaN operands (which also raise invalid on x86) are not
// modelled.
// ---------------------------------------------------------------------------
enum { FPE_DIVIDE = 1, FPE_OVER = 2, FPE_UNDER = 4, FPE_INVALID = 8 };
template <typename T> struct FpLim;
template <> struct FpLim<double> {
    static constexpr int emin = -1022;  // smallest normal 2^emin
    __device__ static double tiny() { return 2.2250738585072014e-308; }
};
template <> struct FpLim<float> {
    static constexpr int emin = -126;
    __device__ static float tiny() { return 1.17549435e-38f; }
};

// m * 2^e with m the correctly rounded mantissa result (m = exact - err):
// tiny after rounding with unbounded exponent, and not exact on the
// subnormal grid
template <typename T> __device__ inline bool fpe_underflow(T m, T err, int e) {
    int em;
    (void)frexp(m, &em);
    if (em + e > FpLim<T>::emin) return false;
    if (err != T(0)) return true;
    return ldexp(ldexp(m, e), -e) != m;
}

template <typename T, int OPK> __device__ inline uint32_t fpe_arith(T a, T b, T r) {
    const T ar = fabs(r);
    if (ar > FpLim<T>::tiny() && ar < T(INFINITY)) return 0;  // a normal result raises nothing
    if (r != r) return (a == a && b == b) ? FPE_INVALID : 0;
    const bool fin = isfinite(a) && isfinite(b);
    if (ar == T(INFINITY)) {
        if (!fin) return 0;
        return (OPK == IONP_TRUE_DIVIDE && b == T(0)) ? FPE_DIVIDE : FPE_OVER;
    }
    if constexpr (OPK == IONP_MULTIPLY || OPK == IONP_TRUE_DIVIDE) {
        if (!fin || a == T(0) || b == T(0)) return 0;
        int ea, eb;
        const T ma = frexp(a, &ea), mb = frexp(b, &eb);  // |m| in [0.5, 1)
        if constexpr (OPK == IONP_MULTIPLY) {
            const T m = ma * mb;
            return fpe_underflow(m, fma(ma, mb, -m), ea + eb) ? FPE_UNDER : 0;
        } else {
            const T m = ma / mb;
            return fpe_underflow(m, fma(-m, mb, ma), ea - eb) ? FPE_UNDER : 0;
        }
    }
    return 0;
}
// dynamic-op form (generic kernel): arithmetic ops only
template <typename T> __device__ inline uint32_t fpe_arith_dyn(int op, T a, T b, T r) {
    switch (op) {
        case IONP_ADD: return fpe_arith<T, IONP_ADD>(a, b, r);
        case IONP_SUBTRACT: return fpe_arith<T, IONP_SUBTRACT>(a, b, r);
        case IONP_MULTIPLY: return fpe_arith<T, IONP_MULTIPLY>(a, b, r);
        case IONP_TRUE_DIVIDE: return fpe_arith<T, IONP_TRUE_DIVIDE>(a, b, r);
        default: return 0;
    }
}
// casts from a float dtype (NumPy's C conversions on x86-64): to an
// integer, invalid when the value does not fit the hardware conversion the
// compiler uses (32-bit for int8..int32, 64-bit for uint32 / int64, the
// two-step 64-bit sequence for uint64); float64 -> float32 overflows or
// underflows like an arithmetic result
// the flags of casting the float32 / float64 value x (as double) to out_dt
__device__ inline uint32_t fpe_cast_x(int in_dt, int out_dt, double x) {
    switch (out_dt) {
        case IONP_INT8:
        case IONP_UINT8:
        case IONP_INT16:
        case IONP_UINT16:
        case IONP_INT32: return (x > -2147483649.0 && x < 2147483648.0) ? 0 : FPE_INVALID;
        case IONP_UINT32:
        case IONP_INT64: return (x >= -9223372036854775808.0 && x < 9223372036854775808.0) ? 0 : FPE_INVALID;
        case IONP_UINT64: return (x >= -9223372036854775808.0 && x < 18446744073709551616.0) ? 0 : FPE_INVALID;
        case IONP_FLOAT32: {
            if (in_dt != IONP_FLOAT64 || !isfinite(x)) return 0;
            const float f = (float)x;
            if (isinf(f)) return FPE_OVER;
            if (fabsf(f) > FpLim<float>::tiny() || x == 0.0) return 0;
            int ex, em;
            const double mx = frexp(x, &ex);
            (void)frexpf((float)mx, &em);  // the 24-bit rounding, unbounded exponent
            return (em + ex <= FpLim<float>::emin && (double)f != x) ? FPE_UNDER : 0;
        }
        default: return 0;
    }
}
__device__ inline uint32_t fpe_cast(int in_dt, int out_dt, const char *p) {
    if (!is_float_dt(in_dt)) return 0;
    return fpe_cast_x(in_dt, out_dt, in_dt == IONP_FLOAT32 ? (double)*(const float *)p : *(const double *)p);
}
// OR of the wavefront's flags into *word by one lane (all lanes call it)
__device__ inline void fpe_commit(uint32_t *word, uint32_t fl) {
    uint32_t w = 0;
    if (__ballot(fl & FPE_DIVIDE)) w |= FPE_DIVIDE;
    if (__ballot(fl & FPE_OVER)) w |= FPE_OVER;
    if (__ballot(fl & FPE_UNDER)) w |= FPE_UNDER;
    if (__ballot(fl & FPE_INVALID)) w |= FPE_INVALID;
    if (w && (threadIdx.x & 63) == 0) atomicOr(word, w);
}

// One output element of the generic kernel with its IEEE flags.
template <int LT, class S> __device__ inline uint32_t ew_element_fpe_src(const EwArgs &A, const S &s, char *po) {
    using T = typename dt_type<LT>::T;
    if (A.op == OP_CAST) {
        store_as<T>(A.out_dt, po, s.template val<LT>(0, A.in_dt[0]));
        return is_float_dt(A.in_dt[0]) ? fpe_cast_x(A.in_dt[0], A.out_dt, s.fx(0, A.in_dt[0])) : 0;
    }
    if constexpr (std::is_floating_point<T>::value) {
        if (A.op <= IONP_TRUE_DIVIDE) {
            const T a = s.template val<LT>(0, A.in_dt[0]), b = s.template val<LT>(1, A.in_dt[1]);
            const T r = arith_op<LT>(A.op, a, b);
            store_as<T>(A.out_dt, po, r);
            return fpe_arith_dyn<T>(A.op, a, b, r);
        }
    }
    ew_element_src<LT>(A, s, po);
    return 0;
}
template <int LT>
__device__ inline uint32_t ew_element_fpe(const EwArgs &A, const char *p0, const char *p1, const char *p2, char *po) {
    return ew_element_fpe_src<LT>(A, PtrSrc{{p0, p1, p2}}, po);
}

// ---------------------------------------------------------------------------
// float16 elements (NumPy's HALF loops): a float16 loop computes in float32
// and rounds the result to float16 (npy_float_to_half); maximum / minimum
// return an operand's own bits, the first on ties and when it is NaN
// (npy_half_ge / npy_half_le are false on NaN); comparisons compare the
// exact float values.  Casts to float16 round the float32 / float64 source
// directly (integers through float64, exact in float16's range); casts from
// float16 go through the exact float64 value.
// ---------------------------------------------------------------------------
// any dtype as float16 bits (the buffered cast NumPy applies first)
__device__ inline uint16_t to_half_bits(int dt, const char *p) {
    uint32_t fl = 0;
    switch (dt) {
        case IONP_FLOAT16: return *(const uint16_t *)p;
        case IONP_FLOAT32: return f32_to_f16_bits(*(const uint32_t *)p, &fl);
        case IONP_FLOAT64: return f64_to_f16_bits(*(const uint64_t *)p, &fl);
        default: return d2h(load_as<IONP_FLOAT64>(dt, p), &fl);
    }
}
// one element-wise float16 step acc (op) x: add / multiply in float32
// rounded back, maximum / minimum keep acc on ties and when it is NaN
__device__ inline uint16_t hstep(int op, uint16_t acc, uint16_t x) {
    const float a = h2f(acc), b = h2f(x);
    uint32_t fl = 0;
    switch (op) {
        case IONP_ADD: return f2h(fadd(a, b), &fl);
        case IONP_MULTIPLY: return f2h(fmul(a, b), &fl);
        case IONP_MAXIMUM: return (a >= b || a != a) ? acc : x;
        default: return (a <= b || a != a) ? acc : x;
    }
}
// value of a float16-loop operand (float16, bool or an 8-bit integer): exact
__device__ inline float hval(int dt, const char *p) {
    return dt == IONP_FLOAT16 ? h2f(*(const uint16_t *)p) : load_as<IONP_FLOAT32>(dt, p);
}
// the same operand as float16 bits (exact)
__device__ inline uint16_t hbits(int dt, const char *p) { return to_half_bits(dt, p); }
// any dtype in loop dtype LT (float32 / float64), float16 exactly
template <int LT> __device__ inline typename dt_type<LT>::T load_hx(int dt, const char *p) {
    if (dt == IONP_FLOAT16) {
        const uint16_t h = *(const uint16_t *)p;
        if constexpr (LT == IONP_FLOAT32) return h2f(h);
        else return cvt<LT>(h2d(h));
    }
    return load_as<LT>(dt, p);
}
// element cast with float16 on either side; returns its IEEE flags
__device__ inline uint32_t cast_half(int in_dt, const char *p, int out_dt, char *po) {
    uint32_t fl = 0;
    if (out_dt == IONP_FLOAT16) {
        uint16_t h;
        switch (in_dt) {
            case IONP_FLOAT16: h = *(const uint16_t *)p; break;
            case IONP_FLOAT32: h = f32_to_f16_bits(*(const uint32_t *)p, &fl); break;
            case IONP_FLOAT64: h = f64_to_f16_bits(*(const uint64_t *)p, &fl); break;
            default: h = d2h(load_as<IONP_FLOAT64>(in_dt, p), &fl); break;
        }
        *(uint16_t *)po = h;
        return fl;
    }
    if (in_dt != IONP_FLOAT16) {  // neither side float16 (a where operand)
        store_as<double>(out_dt, po, load_as<IONP_FLOAT64>(in_dt, p));
        return 0;
    }
    const uint16_t h = *(const uint16_t *)p;
    if (out_dt == IONP_FLOAT32) {
        *(uint32_t *)po = f16_to_f32_bits(h);
        return 0;
    }
    if (out_dt == IONP_FLOAT64) {
        *(uint64_t *)po = f16_to_f64_bits(h);
        return 0;
    }
    const double x = h2d(h);
    store_as<double>(out_dt, po, x);
    return fpe_cast(IONP_FLOAT64, out_dt, (const char *)&x);
}
// one element of an elementwise call that involves float16
__device__ inline uint32_t ew_half_elem(const EwArgs &A, const char *p0, const char *p1, const char *p2, char *po) {
    const int op = A.op;
    if (op == OP_CAST) return cast_half(A.in_dt[0], p0, A.out_dt, po);
    if (op == OP_WHERE) {
        const bool c = load_as<IONP_BOOL>(A.in_dt[0], p0) != 0;
        return cast_half(c ? A.in_dt[1] : A.in_dt[2], c ? p1 : p2, A.out_dt, po);
    }
    if (A.loop_dt == IONP_FLOAT16) {
        const float a = hval(A.in_dt[0], p0), b = hval(A.in_dt[1], p1);
        if (op >= IONP_LESS) {
            *(uint8_t *)po = (uint8_t)cmp_op<float>(op, a, b);
            return 0;
        }
        if (op == IONP_MAXIMUM || op == IONP_MINIMUM) {
            const bool first = (op == IONP_MAXIMUM ? a >= b : a <= b) || a != a;
            *(uint16_t *)po = first ? hbits(A.in_dt[0], p0) : hbits(A.in_dt[1], p1);
            return 0;
        }
        const float r = arith_op<IONP_FLOAT32>(op, a, b);
        uint32_t fl = fpe_arith_dyn<float>(op, a, b, r);
        *(uint16_t *)po = f2h(r, &fl);
        return fl;
    }
    if (A.loop_dt == IONP_FLOAT32) {
        const float a = load_hx<IONP_FLOAT32>(A.in_dt[0], p0), b = load_hx<IONP_FLOAT32>(A.in_dt[1], p1);
        if (op >= IONP_LESS) {
            *(uint8_t *)po = (uint8_t)cmp_op<float>(op, a, b);
            return 0;
        }
        const float r = arith_op<IONP_FLOAT32>(op, a, b);
        store_as<float>(A.out_dt, po, r);
        return fpe_arith_dyn<float>(op, a, b, r);
    }
    const double a = load_hx<IONP_FLOAT64>(A.in_dt[0], p0), b = load_hx<IONP_FLOAT64>(A.in_dt[1], p1);
    if (op >= IONP_LESS) {
        *(uint8_t *)po = (uint8_t)cmp_op<double>(op, a, b);
        return 0;
    }
    const double r = arith_op<IONP_FLOAT64>(op, a, b);
    store_as<double>(A.out_dt, po, r);
    return fpe_arith_dyn<double>(op, a, b, r);
}

// ---------------------------------------------------------------------------
// index helpers
// ---------------------------------------------------------------------------
// decompose idx over dims (innermost first) and dot with strides
__device__ inline int64_t dot_idx(uint32_t idx, int n, const int64_t *shape, const FastDiv *dv, const int64_t *st) {
    int64_t off = 0;
    for (int d = 0; d < n; d++) {
        uint32_t c;
        if (d + 1 < n) {
            const uint32_t q = dv[d].div(idx);
            c = idx - q * (uint32_t)shape[d];
            idx = q;
        } else {
            c = idx;
        }
        off += (int64_t)c * st[d];
    }
    return off;
}
__device__ inline void out_offsets(const RedArgs &A, uint32_t o, int64_t *in_off, int64_t *out_off) {
    int64_t a = 0, b = 0;
    for (int d = 0; d < A.nk; d++) {
        uint32_t c;
        if (d + 1 < A.nk) {
            const uint32_t q = A.k_div[d].div(o);
            c = o - q * (uint32_t)A.k_shape[d];
            o = q;
        } else {
            c = o;
        }
        a += (int64_t)c * A.k_in_st[d];
        b += (int64_t)c * A.k_out_st[d];
    }
    *in_off = a;
    *out_off = b;
}
__device__ inline int64_t w_offset(const RedArgs &A, int64_t q) {
    if (A.nw == 1) return q * A.w_st[0];
    return dot_idx((uint32_t)q, A.nw, A.w_shape, A.w_div, A.w_st);
}
__device__ inline int64_t r_offset(const RedArgs &A, int64_t r) {
    if (A.nr == 0) return 0;
    if (A.nr == 1) return r * A.r_st[0];
    return dot_idx((uint32_t)r, A.nr, A.r_shape, A.r_div, A.r_st);
}
// where= mask offsets (bool bytes) along the same kept / W / outer axes
__device__ inline int64_t mask_k_offset(const RedArgs &A, uint32_t o) {
    return dot_idx(o, A.nk, A.k_shape, A.k_div, A.k_m_st);
}
__device__ inline int64_t mask_w_offset(const RedArgs &A, int64_t q) {
    if (A.nw == 1) return q * A.w_m_st[0];
    return dot_idx((uint32_t)q, A.nw, A.w_shape, A.w_div, A.w_m_st);
}
__device__ inline int64_t mask_r_offset(const RedArgs &A, int64_t r) {
    if (A.nr == 0) return 0;
    if (A.nr == 1) return r * A.r_m_st[0];
    return dot_idx((uint32_t)r, A.nr, A.r_shape, A.r_div, A.r_m_st);
}

template <int LT> __device__ inline typename dt_type<LT>::T ld(int dt, const char *base, int64_t off) {
    switch (dt) {
        case IONP_BOOL: return cvt<LT>((uint8_t)(((const uint8_t *)base)[off] != 0));
        case IONP_INT8: return cvt<LT>(((const int8_t *)base)[off]);
        case IONP_UINT8: return cvt<LT>(((const uint8_t *)base)[off]);
        case IONP_INT16: return cvt<LT>(((const int16_t *)base)[off]);
        case IONP_UINT16: return cvt<LT>(((const uint16_t *)base)[off]);
        case IONP_INT32: return cvt<LT>(((const int32_t *)base)[off]);
        case IONP_UINT32: return cvt<LT>(((const uint32_t *)base)[off]);
        case IONP_INT64: return cvt<LT>(((const int64_t *)base)[off]);
        case IONP_UINT64: return cvt<LT>(((const uint64_t *)base)[off]);
        case IONP_FLOAT32: return cvt<LT>(((const float *)base)[off]);
        case IONP_FLOAT16: return cvt<LT>(h2f(((const uint16_t *)base)[off]));  // exact in float32
        default: return cvt<LT>(((const double *)base)[off]);
    }
}

// Integer / bool reduce result in the output dtype: wrapped to its width;
// add on bool is logical or (bool_same, ufunc.rs:2341).
template <int LT, typename AT> __device__ inline typename dt_type<LT>::T int_result(AT acc) {
    using ST = typename dt_type<LT>::T;
    if constexpr (LT == IONP_BOOL) return (ST)(acc != 0);
    else return (ST)acc;
}

// Element `off` of a buffer of input type IT, cast to loop dtype LT (the
// conversion ld<LT> applies for the matching dtype code).
template <int LT, typename IT> __device__ inline typename dt_type<LT>::T ld_t(const char *base, int64_t off) {
    if constexpr (std::is_same<IT, bool>::value) return cvt<LT>((uint8_t)(((const uint8_t *)base)[off] != 0));
    else return cvt<LT>(((const IT *)base)[off]);
}

// ld_t through a streaming (non-temporal) load
template <int LT, typename IT> __device__ inline typename dt_type<LT>::T ld_t_nt(const char *base, int64_t off) {
    if constexpr (std::is_same<IT, bool>::value)
        return cvt<LT>((uint8_t)(__builtin_nontemporal_load((const uint8_t *)base + off) != 0));
    else
        return cvt<LT>(__builtin_nontemporal_load((const IT *)base + off));
}

// f(IT()) for the element type of dtype code dt (bool for IONP_BOOL)
template <typename F> __device__ inline auto in_type_dispatch(int dt, const F &f) {
    switch (dt) {
        case IONP_BOOL: return f(bool());
        case IONP_INT8: return f(int8_t());
        case IONP_UINT8: return f(uint8_t());
        case IONP_INT16: return f(int16_t());
        case IONP_UINT16: return f(uint16_t());
        case IONP_INT32: return f(int32_t());
        case IONP_UINT32: return f(uint32_t());
        case IONP_INT64: return f(int64_t());
        case IONP_UINT64: return f(uint64_t());
        case IONP_FLOAT32: return f(float());
        default: return f(double());
    }
}

// Starting value of a float add fold: initial= when given, else 0.
template <typename T> __device__ inline T fseed(const RedArgs &A) {
    if (!A.has_init) return T(0);
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, A.init_bits);
    else return __builtin_bit_cast(T, (uint32_t)A.init_bits);
}

template <typename T> struct FDT;
template <> struct FDT<float> { static constexpr int v = IONP_FLOAT32; };
template <> struct FDT<double> { static constexpr int v = IONP_FLOAT64; };

// ---------------------------------------------------------------------------
// Narrow float add (W empty): acc = 0; acc = acc + x_r for the reduced
// elements in iterator order, starting at input offset ib.  The loads do not
// depend on acc, so U of them are issued ahead of the dependent add chain.
// The chain uses plain IEEE adds: they equal fadd whenever no NaN arises,
// and a NaN, once produced, stays NaN to the end; an output that ends NaN
// is recomputed with fadd's x86 NaN rules.
// ---------------------------------------------------------------------------
template <typename T, typename IT, int U> __device__ inline T narrow_chain(const RedArgs &A, int64_t ib) {
    constexpr int LT = FDT<T>::v;
    T acc = fseed<T>(A);
    int64_t r = 0;
    if (A.nr == 1) {
        const int64_t st = A.r_st[0];
        const char *x = A.in;
        T cur[U], nxt[U];
        if (A.rsize >= U) {
#pragma unroll
            for (int u = 0; u < U; u++) cur[u] = ld_t_nt<LT, IT>(x, ib + (int64_t)u * st);
            r = U;
            for (; r + U <= A.rsize; r += U) {
#pragma unroll
                for (int u = 0; u < U; u++) nxt[u] = ld_t_nt<LT, IT>(x, ib + (r + u) * st);
#pragma unroll
                for (int u = 0; u < U; u++) acc = acc + cur[u];
#pragma unroll
                for (int u = 0; u < U; u++) cur[u] = nxt[u];
            }
#pragma unroll
            for (int u = 0; u < U; u++) acc = acc + cur[u];
        }
        for (; r < A.rsize; r++) acc = acc + ld_t<LT, IT>(x, ib + r * st);
        if (acc != acc) {
            acc = fseed<T>(A);
            for (int64_t q = 0; q < A.rsize; q++) acc = fadd(acc, ld_t<LT, IT>(x, ib + q * st));
        }
        return acc;
    }
    constexpr int V = 8;
    for (; r + V <= A.rsize; r += V) {
        T v[V];
#pragma unroll
        for (int u = 0; u < V; u++) v[u] = ld_t<LT, IT>(A.in, ib + r_offset(A, r + u));
#pragma unroll
        for (int u = 0; u < V; u++) acc = acc + v[u];
    }
    for (; r < A.rsize; r++) acc = acc + ld_t<LT, IT>(A.in, ib + r_offset(A, r));
    if (acc != acc) {
        acc = fseed<T>(A);
        for (int64_t q = 0; q < A.rsize; q++) acc = fadd(acc, ld_t<LT, IT>(A.in, ib + r_offset(A, q)));
    }
    return acc;
}
// U: loads in flight per thread (registers: 2 U values)
template <typename T, int U = (sizeof(T) == 8 ? 24 : 32)> __device__ inline T narrow_sum(const RedArgs &A, int64_t ib) {
    return in_type_dispatch(A.in_dt, [&](auto it) { return narrow_chain<T, decltype(it), U>(A, ib); });
}

// Value of v in lane `lane` (uniform) of the wavefront.
template <typename T> __device__ inline T readlane_t(T v, int lane) {
    if constexpr (sizeof(T) == 8) {
        uint64_t b;
        std::memcpy(&b, &v, 8);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
        const uint64_t r = ((uint64_t)hi << 32) | lo;
        T out;
        std::memcpy(&out, &r, 8);
        return out;
    } else if constexpr (sizeof(T) == 4) {
        uint32_t b;
        std::memcpy(&b, &v, 4);
        const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)b, lane);
        T out;
        std::memcpy(&out, &r, 4);
        return out;
    } else {  // 1- and 2-byte integers
        return (T)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    }
}

// One wavefront runs one output's chain acc = seed; acc = acc + x[r] (r in
// iterator order).  Blocks of 64 B elements are loaded one per lane and slot
// (the next block's loads in flight meanwhile), staged in the wavefront's
// own LDS region `lb` in element order, and added from
// broadcast 16-byte LDS reads: the chain then runs at the dependent-add
// latency (readlane operands, one per element, ran about 3 times slower on
// MI355X).  Plain IEEE adds agree with fadd until a NaN arises, and a NaN
// stays, so an output that ends NaN is redone with fadd's x86 NaN rules.
// R: 16-byte LDS reads per chunk (registers: 2 R vectors); B: loads per lane
// per block (a block is 64 B elements, `lb` holds one block)
template <typename T, typename IT, int R = 8, int B = 4>
__device__ inline T narrow_wave_chain(const RedArgs &A, int64_t ib, T *lb) {
    constexpr int LT = FDT<T>::v;
    constexpr int BLK = 64 * B;
    constexpr int CH = 16 / sizeof(T);
    using V = T __attribute__((ext_vector_type(CH)));
    const int lane = threadIdx.x & 63;
    const int64_t n = A.rsize;
    auto off = [&](int64_t r) { return A.nr == 1 ? r * A.r_st[0] : r_offset(A, r); };
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    T cur[B], nxt[B];
#pragma unroll
    for (int k = 0; k < B; k++) {
        const int64_t r = k * 64 + lane;
        cur[k] = r < n ? ld_t<LT, IT>(A.in, ib + off(r)) : T(0);
    }
    T acc = fseed<T>(A);
    for (int64_t rb = 0; rb < n; rb += BLK) {
        const int64_t nb = rb + BLK;
        if (nb < n) {
#pragma unroll
            for (int k = 0; k < B; k++) {
                const int64_t r = nb + k * 64 + lane;
                nxt[k] = r < n ? ld_t<LT, IT>(A.in, ib + off(r)) : T(0);
            }
        }
#pragma unroll
        for (int k = 0; k < B; k++) lb[k * 64 + lane] = cur[k];
        wave_sync();
        if (nb <= n) {
            // chunks of R reads, the next chunk's reads issued before the
            // current chunk's adds (kept there by the scheduling barriers),
            // so LDS latency stays off the chain
            constexpr int CK = R * CH, NCK = BLK / CK;
            V ch[2][R];
#pragma unroll
            for (int j = 0; j < R; j++) ch[0][j] = *(const V *)(lb + j * CH);
#pragma unroll
            for (int c = 0; c < NCK; c++) {
                if (c + 1 < NCK) {
#pragma unroll
                    for (int j = 0; j < R; j++) ch[(c + 1) & 1][j] = *(const V *)(lb + (c + 1) * CK + j * CH);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < R; j++)
#pragma unroll
                    for (int e = 0; e < CH; e++) acc = acc + ch[c & 1][j][e];
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
            const int cnt = (int)(n - rb);
            for (int e = 0; e < cnt; e++) acc = acc + lb[e];
        }
        wave_sync();  // every lane has read the block before the next one is stored
#pragma unroll
        for (int k = 0; k < B; k++) cur[k] = nxt[k];
    }
    if (acc != acc) {
        acc = fseed<T>(A);
        for (int64_t q = 0; q < n; q++) acc = fadd(acc, ld_t<LT, IT>(A.in, ib + off(q)));
    }
    return acc;
}
// lb: the wavefront's LDS region of 64 B elements, 16-byte aligned
template <typename T, int R = 8, int B = 4> __device__ inline T narrow_wave_sum(const RedArgs &A, int64_t ib, T *lb) {
    return in_type_dispatch(A.in_dt, [&](auto it) { return narrow_wave_chain<T, decltype(it), R, B>(A, ib, lb); });
}

// ---------------------------------------------------------------------------
// NumPy's pairwise tree over an LDS window (ufunc.rs:6416-6448)
// ---------------------------------------------------------------------------
// Node split of pairwise_sum: n2 = n/2 rounded down to a multiple of 8.
__device__ inline int pw_split(int n) { return (n >> 1) & ~7; }

// pairwise_sum of x[0..n) for n <= 128 (a single leaf) in one thread: n < 8
// adds in order from -0.0; otherwise 8 interleaved accumulators combined as
// ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the n % 8 tail in order.
template <typename T, typename Get> __device__ inline T pw_leaf(const Get &get, int n) {
    if (n < 8) {
        T r = T(-0.0);  // NumPy 2.2.6's short leaf: a sum of -0.0 stays -0.0
        for (int i = 0; i < n; i++) r = fadd(r, get(i));
        return r;
    }
    T r[8];
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = get(j);
    int i = 8;
    for (; i + 8 <= n; i += 8) {
        T x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = get(i + j);
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = fadd(r[j], x[j]);
    }
    T res = fadd(fadd(fadd(r[0], r[1]), fadd(r[2], r[3])), fadd(fadd(r[4], r[5]), fadd(r[6], r[7])));
    for (; i < n; i++) res = fadd(res, get(i));
    return res;
}

constexpr int kPwMaxWin = 8192;
constexpr int kPwThreads = 512;

// Evaluates pairwise(x[0..L)) with the whole block, x[i] = get(i) (an LDS
// copy or a global gather); returns the result in every thread.  Scratch:
// racc (8 per leaf), val (per octet), leaf table.
template <typename T, typename Get>
__device__ T pairwise_block(const Get &get, int L, T *racc, T *val, int *leaf_lo, int *leaf_n, int *cnt) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int noct = (L + 7) >> 3;
    // 1. leaf starts: a leaf begins at octet o iff the descent for element
    //    8o ends at a leaf whose first element is 8o (every leaf starts on a
    //    multiple of 8).  The table order is irrelevant: chains address
    //    leaves through it and leaf values are stored by first octet.
    if (tid == 0) cnt[0] = 0;
    __syncthreads();
    for (int o = tid; o < noct; o += nt) {
        int lo = 0, n = L;
        while (n > 128) {
            const int n2 = pw_split(n);
            if (8 * o < lo + n2) n = n2;
            else { lo += n2; n -= n2; }
        }
        if (lo == 8 * o) {
            const int idx = atomicAdd(&cnt[0], 1);
            leaf_lo[idx] = lo;
            leaf_n[idx] = n;
        }
    }
    __syncthreads();
    const int nleaves = cnt[0];
    // 2. the 8 accumulator chains of every leaf, each in element order
    for (int c = tid; c < 8 * nleaves; c += nt) {
        const int l = c >> 3, j = c & 7, lo = leaf_lo[l], n = leaf_n[l];
        if (n < 8) {
            if (j == 0) {
                T r = T(-0.0);  // short leaf: starts from -0.0 as NumPy's does
                for (int i = 0; i < n; i++) r = fadd(r, get(lo + i));
                racc[c] = r;
            }
        } else {
            T r = get(lo + j);
            const int m = n >> 3;
            for (int k = 1; k < m; k++) r = fadd(r, get(lo + 8 * k + j));
            racc[c] = r;
        }
    }
    __syncthreads();
    // 3. leaf results: ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)), then the n%8 tail
    for (int l = tid; l < nleaves; l += nt) {
        const int lo = leaf_lo[l], n = leaf_n[l];
        const T *r = racc + 8 * l;
        T res;
        if (n < 8) {
            res = r[0];
        } else {
            res = fadd(fadd(fadd(r[0], r[1]), fadd(r[2], r[3])), fadd(fadd(r[4], r[5]), fadd(r[6], r[7])));
            for (int i = lo + (n & ~7); i < lo + n; i++) res = fadd(res, get(i));
        }
        val[lo >> 3] = res;
    }
    __syncthreads();
    // 4. internal nodes, deepest level first: node value = left + right,
    //    stored at the node's first octet.
    for (int d = 9; d >= 0; d--) {  // leaves hold >= 64 elements: depth <= 8 for L <= 8192
        if (tid < (1 << d)) {
            int lo = 0, n = L;
            bool ok = true;
            for (int lev = 0; lev < d; lev++) {
                if (n <= 128) { ok = false; break; }
                const int n2 = pw_split(n);
                if ((tid >> (d - 1 - lev)) & 1) { lo += n2; n -= n2; }
                else n = n2;
            }
            if (ok && n > 128) {
                const int n2 = pw_split(n);
                val[lo >> 3] = fadd(val[lo >> 3], val[(lo + n2) >> 3]);
            }
        }
        __syncthreads();
    }
    return val[0];
}


}  // namespace ionp
