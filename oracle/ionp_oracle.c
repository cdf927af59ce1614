/*
 * ionp_oracle.c -- CPU restatement of anionpy's ufunc / reduction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP backend:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (anionpy_amd) never links, imports or calls it.
 *
 * It restates, in plain scalar C, what the reference computes and what its
 * own oracle (NumPy 2.2.6 on x86-64, the library the reference's
 * differential suite compares against byte-for-byte) returns:
 *
 *   - dtype promotion            ionp-core/src/dtype.rs:297-357
 *   - binary loop dtype          ionp-core/src/ufunc.rs:1101-1196
 *   - scalar kernels             ionp-core/src/ufunc.rs:2305-2429 (int_same,
 *                                float_max/min), 2460-2579 (float_same),
 *                                3479 (cmp_same)
 *   - broadcast + NdIter walk    ionp-core/src/shape.rs:54-146, 317-387
 *   - where_select               ionp-core/src/sort.rs:1869-1914
 *   - casts                      ionp-core/src/buffer.rs:81-210
 *   - pairwise add.reduce        ionp-core/src/ufunc.rs:6349-6448 (tree),
 *                                6481-6531 (runs), 6658-6799 (narrow/wide),
 *                                7091-7167 (dispatch, 0 seed)
 *   - maximum/minimum.reduce     ionp-core/src/ufunc.rs:2402-2429, 6546
 *   - argmax/argmin              ionp-core/src/ufunc.rs:8037-8206
 *
 * Where the reference and its NumPy oracle differ, the oracle follows NumPy
 * (the reference's own contract is "match NumPy"):
 *   (1) add.reduce is evaluated per NumPy iterator buffer window of
 *       `bufsize` (8192) elements: acc = 0; acc = acc + pw(window) ...
 *   (2) maximum/minimum ties (+0 vs -0) return the second operand; NaN
 *       results follow the x86 AVX-512 reduce loop (canonical quiet NaN for
 *       NaNs seen by the vector part, the NaN's own bits otherwise).
 *   (3) NaN results of + - * / follow x86 SSE rules: first NaN operand
 *       (quieted), else the default NaN (sign bit set).
 *   (4) float -> integer casts out of range follow x86 cvtt* results.
 *
 * Every function here is a direct, unoptimised restatement; it is built with
 * -O2 -ffp-contract=off and no fast-math so that each IEEE operation is one
 * correctly-rounded operation in program order.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <fenv.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_MAXDIM 16

enum {
    DT_BOOL = 0, DT_I8, DT_U8, DT_I16, DT_U16, DT_I32, DT_U32, DT_I64, DT_U64,
    DT_F32, DT_F64, DT_F16, DT_COUNT
};
static const int ITEMSIZE[DT_COUNT] = {1, 1, 1, 2, 2, 4, 4, 8, 8, 4, 8, 2};

enum {
    OP_ADD = 0, OP_SUBTRACT, OP_MULTIPLY, OP_TRUE_DIVIDE, OP_MAXIMUM, OP_MINIMUM,
    OP_LESS, OP_LESS_EQUAL, OP_GREATER, OP_GREATER_EQUAL, OP_EQUAL, OP_NOT_EQUAL,
    OP_COUNT
};

/* status codes: 0 ok; the rest mirror IonpError (ionp-core/src/error.rs:8-42) */
enum { ORA_OK = 0, ORA_EBROADCAST = 1, ORA_ETYPE = 3, ORA_EVALUE = 4, ORA_ENOLOOP = 5, ORA_EAXIS = 6 };

typedef struct {
    int32_t dtype;
    int32_t ndim;
    int64_t shape[ORA_MAXDIM];
    int64_t strides[ORA_MAXDIM]; /* in elements, may be 0 or negative */
    void *data;                  /* address of element (0, ..., 0) */
} ora_desc;

/* ------------------------------------------------------------------------ */
/* dtype classification and promotion (dtype.rs:297-357)                    */
/* ------------------------------------------------------------------------ */
static int is_float(int dt) { return dt == DT_F32 || dt == DT_F64; }
static int is_float_any(int dt) { return is_float(dt) || dt == DT_F16; }
static int is_signed_int(int dt) { return dt == DT_I8 || dt == DT_I16 || dt == DT_I32 || dt == DT_I64; }
static int is_unsigned_int(int dt) { return dt == DT_U8 || dt == DT_U16 || dt == DT_U32 || dt == DT_U64; }
static int signed_of_size(int bytes)
{
    switch (bytes) { case 1: return DT_I8; case 2: return DT_I16; case 4: return DT_I32; default: return DT_I64; }
}
static int unsigned_of_size(int bytes)
{
    switch (bytes) { case 1: return DT_U8; case 2: return DT_U16; case 4: return DT_U32; default: return DT_U64; }
}

/* NumPy's result_type over {bool, int8..uint64, float32, float64}. */
int ora_result_type(int a, int b)
{
    if (a == b) return a;
    if (a == DT_BOOL) return b;
    if (b == DT_BOOL) return a;
    if (a == DT_F16 || b == DT_F16) { /* float16 holds the 8-bit integers exactly */
        int o = a == DT_F16 ? b : a;
        if (ITEMSIZE[o] == 1) return DT_F16;
        if (o == DT_F32 || ITEMSIZE[o] == 2) return DT_F32;
        return DT_F64;
    }
    if (is_float(a) && is_float(b)) return (a == DT_F64 || b == DT_F64) ? DT_F64 : DT_F32;
    if (is_float(a) || is_float(b)) {
        int f = is_float(a) ? a : b, i = is_float(a) ? b : a;
        if (f == DT_F64) return DT_F64;
        return ITEMSIZE[i] <= 2 ? DT_F32 : DT_F64; /* float32 holds int16 exactly, not int32 */
    }
    if (is_signed_int(a) == is_signed_int(b)) {
        int sz = ITEMSIZE[a] > ITEMSIZE[b] ? ITEMSIZE[a] : ITEMSIZE[b];
        return is_signed_int(a) ? signed_of_size(sz) : unsigned_of_size(sz);
    }
    {
        int s = is_signed_int(a) ? a : b, u = is_signed_int(a) ? b : a;
        if (ITEMSIZE[u] < ITEMSIZE[s]) return s;
        if (ITEMSIZE[u] >= 8) return DT_F64; /* int64 with uint64 */
        return signed_of_size(ITEMSIZE[u] * 2);
    }
}

/* binary_out_dtype (ufunc.rs:1101): loop dtype and output dtype of a binary
 * ufunc.  Returns ORA_OK or an error code. */
int ora_binary_types(int op, int a, int b, int *loop, int *out)
{
    int rt = ora_result_type(a, b);
    if (op == OP_TRUE_DIVIDE) {
        if (!is_float_any(rt)) rt = DT_F64;
        *loop = rt; *out = rt; return ORA_OK;
    }
    if (op == OP_SUBTRACT && rt == DT_BOOL) return ORA_ETYPE;
    if (op >= OP_LESS) {
        *loop = rt; *out = DT_BOOL; return ORA_OK;
    }
    *loop = rt; *out = rt; return ORA_OK;
}

/* reduce_compute_dtype (ufunc.rs:2041-2066): add.reduce of small ints and
 * bool accumulates in the platform long; maximum/minimum keep the dtype. */
int ora_reduce_dtype(int op, int in)
{
    if (op == OP_ADD || op == OP_MULTIPLY) {
        if (in == DT_BOOL || in == DT_I8 || in == DT_I16 || in == DT_I32) return DT_I64;
        if (in == DT_U8 || in == DT_U16 || in == DT_U32) return DT_U64;
    }
    return in;
}

/* ------------------------------------------------------------------------ */
/* scalar value plumbing                                                    */
/* ------------------------------------------------------------------------ */
typedef union { uint64_t u; int64_t i; double d; float f; uint32_t u32; } scal;

static uint64_t f64_bits(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static double bits_f64(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
static uint32_t f32_bits(float x) { uint32_t u; memcpy(&u, &x, 4); return u; }
static float bits_f32(uint32_t u) { float x; memcpy(&x, &u, 4); return x; }

static double quiet64(double x) { return bits_f64(f64_bits(x) | 0x0008000000000000ull); }
static float quiet32(float x) { return bits_f32(f32_bits(x) | 0x00400000u); }
static const uint64_t DEFNAN64 = 0xfff8000000000000ull; /* x86 "real indefinite" */
static const uint32_t DEFNAN32 = 0xffc00000u;
static const uint64_t CANON64 = 0x7ff8000000000000ull;  /* AVX-512 reduce NaN */
static const uint32_t CANON32 = 0x7fc00000u;

/* x86 cvttsd2si semantics: out of range or NaN -> integer indefinite. */
static int32_t cvtt_i32(double x)
{
    if (x != x || x >= 2147483648.0 || x <= -2147483649.0) return INT32_MIN;
    return (int32_t)x;
}
static int64_t cvtt_i64(double x)
{
    if (x != x || x >= 9223372036854775808.0 || x < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}
static uint64_t cvtt_u64(double x)
{
    if (x >= 9223372036854775808.0)
        return (uint64_t)cvtt_i64(x - 9223372036854775808.0) ^ 0x8000000000000000ull;
    return (uint64_t)cvtt_i64(x);
}

/* Load element at p (dtype dt) as one of three carriers. */
static int64_t load_i(int dt, const char *p)
{
    switch (dt) {
    case DT_BOOL: return *(const uint8_t *)p != 0;
    case DT_I8: return *(const int8_t *)p;
    case DT_U8: return *(const uint8_t *)p;
    case DT_I16: return *(const int16_t *)p;
    case DT_U16: return *(const uint16_t *)p;
    case DT_I32: return *(const int32_t *)p;
    case DT_U32: return *(const uint32_t *)p;
    case DT_I64: return *(const int64_t *)p;
    case DT_U64: return (int64_t)*(const uint64_t *)p;
    default: return 0;
    }
}

/* ------------------------------------------------------------------------ */
/* float16: NumPy's npy_half conversions, restated on exact values: a finite */
/* value rounds to nearest even on float16's grid (2^-24 below 2^-14, 11    */
/* significant bits above); |x| >= 65520 overflows to inf; underflow is     */
/* raised when the value is below 2^-14 and not on the grid (tininess       */
/* before rounding).  NaN keeps its top 10 payload bits unquieted (an       */
/* all-zero payload becomes 1).  The flags go to the host status register   */
/* (feraiseexcept), where ora_*_fpe read them like the hardware's own.      */
/* ------------------------------------------------------------------------ */
static uint16_t half_round(double x)
{
    double ax = fabs(x), q, r;
    uint16_t sign = signbit(x) ? 0x8000u : 0;
    int k;
    if (ax == 0) return sign;
    if (isinf(ax)) return (uint16_t)(sign | 0x7c00u); /* exact: no flag */
    if (ax >= 65520.0) {
        feraiseexcept(FE_OVERFLOW);
        return (uint16_t)(sign | 0x7c00u);
    }
    if (ax < 6.103515625e-05) { /* 2^-14: the subnormal grid, units of 2^-24 */
        q = ldexp(ax, 24);
        r = nearbyint(q);
        if (r != q) feraiseexcept(FE_UNDERFLOW);
        return (uint16_t)(sign | (uint16_t)r);
    }
    (void)frexp(ax, &k); /* ax in [2^(k-1), 2^k) */
    q = ldexp(ax, 11 - k); /* in [1024, 2048) */
    r = nearbyint(q);
    if (r == 2048.0) { r = 1024.0; k++; }
    if (k - 1 + 15 >= 31) {
        feraiseexcept(FE_OVERFLOW);
        return (uint16_t)(sign | 0x7c00u);
    }
    return (uint16_t)(sign | ((uint16_t)(k - 1 + 15) << 10) | (uint16_t)((int)r - 1024));
}
static uint16_t f32_to_half(uint32_t fb)
{
    uint32_t a = fb & 0x7fffffffu;
    if (a > 0x7f800000u) { /* NaN */
        uint16_t p = (uint16_t)((a & 0x007fffffu) >> 13);
        return (uint16_t)(((fb >> 16) & 0x8000u) | 0x7c00u | (p ? p : 1));
    }
    return half_round((double)bits_f32(fb));
}
static uint16_t f64_to_half(uint64_t db)
{
    uint64_t a = db & 0x7fffffffffffffffull;
    if (a > 0x7ff0000000000000ull) {
        uint16_t p = (uint16_t)((a & 0x000fffffffffffffull) >> 42);
        return (uint16_t)(((db >> 48) & 0x8000u) | 0x7c00u | (p ? p : 1));
    }
    return half_round(bits_f64(db));
}
/* exact value of a float16 (NaN: its float32 / float64 bits, payload kept) */
static double half_value(uint16_t h)
{
    int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    double v;
    if (e == 31) return bits_f64(((uint64_t)(h & 0x8000u) << 48) | 0x7ff0000000000000ull | ((uint64_t)m << 42));
    v = e == 0 ? ldexp((double)m, -24) : ldexp((double)(1024 + m), e - 25);
    return (h & 0x8000u) ? -v : v;
}
static uint32_t half_f32_bits(uint16_t h)
{
    int e = (h >> 10) & 0x1f;
    if (e == 31) return ((uint32_t)(h & 0x8000u) << 16) | 0x7f800000u | ((uint32_t)(h & 0x3ffu) << 13);
    return f32_bits((float)half_value(h));
}
static float half_f(uint16_t h) { return bits_f32(half_f32_bits(h)); }

/* buffer.rs:81-210 cast matrix, with NumPy's x86 results for NaN payloads
 * and out-of-range float->int. */
static void cast_elem(int sdt, const char *s, int ddt, char *d)
{
    if (sdt == ddt) { memcpy(d, s, ITEMSIZE[sdt]); return; }
    if (ddt == DT_F16) { /* npy_float_to_half / npy_double_to_half; integers through float64 */
        uint16_t h;
        if (sdt == DT_F32) h = f32_to_half(*(const uint32_t *)s);
        else if (sdt == DT_F64) h = f64_to_half(*(const uint64_t *)s);
        else {
            char tmp[8];
            cast_elem(sdt, s, DT_F64, tmp);
            h = half_round(*(const double *)tmp);
        }
        memcpy(d, &h, 2);
        return;
    }
    if (sdt == DT_F16) { /* through the exact float64 value */
        uint16_t h;
        double x;
        memcpy(&h, s, 2);
        if (ddt == DT_F32) { uint32_t b = half_f32_bits(h); memcpy(d, &b, 4); return; }
        x = half_value(h);
        if (ddt == DT_BOOL) { *(uint8_t *)d = (uint8_t)(x != 0.0 || x != x); return; }
        cast_elem(DT_F64, (const char *)&x, ddt, d);
        return;
    }
    if (ddt == DT_BOOL) {
        int nz;
        if (sdt == DT_F64) nz = *(const double *)s != 0.0;
        else if (sdt == DT_F32) nz = *(const float *)s != 0.0f;
        else nz = load_i(sdt, s) != 0;
        *(uint8_t *)d = (uint8_t)nz;
        return;
    }
    if (is_float(sdt)) {
        double x;
        if (sdt == DT_F32) {
            float f = *(const float *)s;
            if (f != f) { /* cvtss2sd: quiet, widen payload */
                uint32_t b = f32_bits(f) | 0x00400000u;
                uint64_t o = ((uint64_t)(b >> 31) << 63) | 0x7ff0000000000000ull | ((uint64_t)(b & 0x7fffff) << 29);
                x = bits_f64(o);
            } else x = (double)f;
        } else x = *(const double *)s;
        switch (ddt) {
        case DT_F64: *(double *)d = x; return;
        case DT_F32:
            if (x != x) { /* cvtsd2ss: quiet, keep the top payload bits */
                uint64_t b = f64_bits(x) | 0x0008000000000000ull;
                uint32_t o = (uint32_t)((b >> 63) << 31) | 0x7f800000u | (uint32_t)((b >> 29) & 0x7fffff);
                *(float *)d = bits_f32(o);
            } else *(float *)d = (float)x;
            return;
        case DT_I8: *(int8_t *)d = (int8_t)(uint8_t)(uint32_t)cvtt_i32(x); return;
        case DT_U8: *(uint8_t *)d = (uint8_t)(uint32_t)cvtt_i32(x); return;
        case DT_I16: *(int16_t *)d = (int16_t)(uint16_t)(uint32_t)cvtt_i32(x); return;
        case DT_U16: *(uint16_t *)d = (uint16_t)(uint32_t)cvtt_i32(x); return;
        case DT_I32: *(int32_t *)d = cvtt_i32(x); return;
        case DT_U32: *(uint32_t *)d = (uint32_t)(uint64_t)cvtt_i64(x); return;
        case DT_I64: *(int64_t *)d = cvtt_i64(x); return;
        case DT_U64: *(uint64_t *)d = cvtt_u64(x); return;
        }
        return;
    }
    /* integer or bool source */
    if (ddt == DT_F64) {
        if (sdt == DT_U64) *(double *)d = (double)*(const uint64_t *)s;
        else *(double *)d = (double)load_i(sdt, s);
        return;
    }
    if (ddt == DT_F32) {
        if (sdt == DT_U64) *(float *)d = (float)*(const uint64_t *)s;
        else *(float *)d = (float)load_i(sdt, s);
        return;
    }
    {
        uint64_t v = (uint64_t)load_i(sdt, s);
        switch (ddt) {
        case DT_I8: case DT_U8: *(uint8_t *)d = (uint8_t)v; return;
        case DT_I16: case DT_U16: *(uint16_t *)d = (uint16_t)v; return;
        case DT_I32: case DT_U32: *(uint32_t *)d = (uint32_t)v; return;
        default: *(uint64_t *)d = v; return;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* scalar kernels: float_same / int_same / cmp_same / float_max / float_min */
/* ------------------------------------------------------------------------ */
static double f64_arith(int op, double a, double b)
{
    double r;
    if (a != a) return quiet64(a);
    if (b != b) return quiet64(b);
    switch (op) {
    case OP_ADD: r = a + b; break;
    case OP_SUBTRACT: r = a - b; break;
    case OP_MULTIPLY: r = a * b; break;
    default: r = a / b; break;
    }
    if (r != r) return bits_f64(DEFNAN64);
    return r;
}
static float f32_arith(int op, float a, float b)
{
    float r;
    if (a != a) return quiet32(a);
    if (b != b) return quiet32(b);
    switch (op) {
    case OP_ADD: r = a + b; break;
    case OP_SUBTRACT: r = a - b; break;
    case OP_MULTIPLY: r = a * b; break;
    default: r = a / b; break;
    }
    if (r != r) return bits_f32(DEFNAN32);
    return r;
}
/* maximum/minimum: NaN in the first operand wins, then NaN in the second,
 * ties (including -0 vs +0) return the second operand; NaNs are not quieted. */
static double f64_max(double a, double b) { if (a != a) return a; if (b != b) return b; return a > b ? a : b; }
static double f64_min(double a, double b) { if (a != a) return a; if (b != b) return b; return a < b ? a : b; }
static float f32_max(float a, float b) { if (a != a) return a; if (b != b) return b; return a > b ? a : b; }
static float f32_min(float a, float b) { if (a != a) return a; if (b != b) return b; return a < b ? a : b; }

static int cmp_d(int op, double a, double b)
{
    switch (op) {
    case OP_LESS: return a < b;
    case OP_LESS_EQUAL: return a <= b;
    case OP_GREATER: return a > b;
    case OP_GREATER_EQUAL: return a >= b;
    case OP_EQUAL: return a == b;
    default: return a != b;
    }
}
static int cmp_s(int op, int64_t a, int64_t b)
{
    switch (op) {
    case OP_LESS: return a < b;
    case OP_LESS_EQUAL: return a <= b;
    case OP_GREATER: return a > b;
    case OP_GREATER_EQUAL: return a >= b;
    case OP_EQUAL: return a == b;
    default: return a != b;
    }
}
static int cmp_u(int op, uint64_t a, uint64_t b)
{
    switch (op) {
    case OP_LESS: return a < b;
    case OP_LESS_EQUAL: return a <= b;
    case OP_GREATER: return a > b;
    case OP_GREATER_EQUAL: return a >= b;
    case OP_EQUAL: return a == b;
    default: return a != b;
    }
}
/* exact int64 vs uint64 comparison (NumPy 2 mixed-sign comparison loops) */
static int cmp_su(int op, int64_t a, uint64_t b)
{
    int lt = a < 0 || (uint64_t)a < b, eq = a >= 0 && (uint64_t)a == b;
    switch (op) {
    case OP_LESS: return lt;
    case OP_LESS_EQUAL: return lt || eq;
    case OP_GREATER: return !lt && !eq;
    case OP_GREATER_EQUAL: return !lt;
    case OP_EQUAL: return eq;
    default: return !eq;
    }
}

static uint64_t wrap_arith(int op, uint64_t a, uint64_t b)
{
    switch (op) {
    case OP_ADD: return a + b;
    case OP_SUBTRACT: return a - b;
    default: return a * b;
    }
}

/* One element of a binary ufunc: operands already cast to the loop dtype.
 * Writes the output element (dtype `odt`). */
static void binary_elem(int op, int ldt, const char *pa, const char *pb, int odt, char *po)
{
    if (ldt == DT_F16) { /* HALF loops: float32 arithmetic rounded back; max/min on the halves */
        uint16_t ha, hb, hr;
        float a, b;
        memcpy(&ha, pa, 2);
        memcpy(&hb, pb, 2);
        a = half_f(ha);
        b = half_f(hb);
        if (op >= OP_LESS) { *(uint8_t *)po = (uint8_t)cmp_d(op, a, b); return; }
        if (op == OP_MAXIMUM) hr = (a != a || (b == b && a >= b)) ? ha : hb;
        else if (op == OP_MINIMUM) hr = (a != a || (b == b && a <= b)) ? ha : hb;
        else hr = f32_to_half(f32_bits(f32_arith(op, a, b)));
        memcpy(po, &hr, 2);
        (void)odt;
        return;
    }
    if (op >= OP_LESS) {
        int r;
        if (ldt == DT_F64) r = cmp_d(op, *(const double *)pa, *(const double *)pb);
        else if (ldt == DT_F32) r = cmp_d(op, *(const float *)pa, *(const float *)pb);
        else if (is_unsigned_int(ldt)) {
            uint64_t a = (uint64_t)load_i(ldt, pa), b = (uint64_t)load_i(ldt, pb);
            if (ldt != DT_U64) { a &= (ITEMSIZE[ldt] == 4 ? 0xffffffffull : ITEMSIZE[ldt] == 2 ? 0xffffull : 0xffull);
                                  b &= (ITEMSIZE[ldt] == 4 ? 0xffffffffull : ITEMSIZE[ldt] == 2 ? 0xffffull : 0xffull); }
            r = cmp_u(op, a, b);
        } else r = cmp_s(op, load_i(ldt, pa), load_i(ldt, pb));
        *(uint8_t *)po = (uint8_t)r;
        (void)odt;
        return;
    }
    if (ldt == DT_F64) {
        double a = *(const double *)pa, b = *(const double *)pb, r;
        if (op == OP_MAXIMUM) r = f64_max(a, b);
        else if (op == OP_MINIMUM) r = f64_min(a, b);
        else r = f64_arith(op, a, b);
        *(double *)po = r;
        return;
    }
    if (ldt == DT_F32) {
        float a = *(const float *)pa, b = *(const float *)pb, r;
        if (op == OP_MAXIMUM) r = f32_max(a, b);
        else if (op == OP_MINIMUM) r = f32_min(a, b);
        else r = f32_arith(op, a, b);
        *(float *)po = r;
        return;
    }
    if (ldt == DT_BOOL) {
        uint8_t a = *(const uint8_t *)pa != 0, b = *(const uint8_t *)pb != 0, r;
        if (op == OP_ADD || op == OP_MAXIMUM) r = a | b;
        else r = a & b; /* multiply, minimum */
        *(uint8_t *)po = r;
        return;
    }
    {
        /* int_same (ufunc.rs:2305): wrapping arithmetic, plain max/min */
        int64_t sa = load_i(ldt, pa), sb = load_i(ldt, pb);
        uint64_t r;
        if (op == OP_MAXIMUM || op == OP_MINIMUM) {
            int gt;
            if (is_unsigned_int(ldt)) gt = (uint64_t)sa > (uint64_t)sb;
            else gt = sa > sb;
            if (op == OP_MAXIMUM) r = (uint64_t)(gt ? sa : sb);
            else r = (uint64_t)(gt ? sb : sa);
        } else r = wrap_arith(op, (uint64_t)sa, (uint64_t)sb);
        switch (ITEMSIZE[ldt]) {
        case 1: *(uint8_t *)po = (uint8_t)r; break;
        case 2: *(uint16_t *)po = (uint16_t)r; break;
        case 4: *(uint32_t *)po = (uint32_t)r; break;
        default: *(uint64_t *)po = r; break;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* broadcasting and the C-order odometer (shape.rs:54-146, 317-387)         */
/* ------------------------------------------------------------------------ */
int ora_broadcast_shapes(int na, const int64_t *a, int nb, const int64_t *b, int *nout, int64_t *out)
{
    int n = na > nb ? na : nb, i;
    for (i = 0; i < n; i++) {
        int64_t da = i < n - na ? 1 : a[i - (n - na)];
        int64_t db = i < n - nb ? 1 : b[i - (n - nb)];
        if (da != db && da != 1 && db != 1) return ORA_EBROADCAST;
        out[i] = da == 1 ? db : da;
    }
    *nout = n;
    return ORA_OK;
}

/* Strides of `d` viewed with the broadcast shape (rank n, shape `shape`). */
static void bstrides(const ora_desc *d, int n, const int64_t *shape, int64_t *st)
{
    int i, off = n - d->ndim;
    for (i = 0; i < n; i++) {
        if (i < off) st[i] = 0;
        else st[i] = (d->shape[i - off] == 1 && shape[i] != 1) ? 0 : d->strides[i - off];
    }
}

/* NumPy nditer axis ordering (npyiter_find_best_axis_ordering), restated
 * for `nop` operands with element strides st[op][axis]: a stable insertion
 * sort from reversed C order, smallest |stride| innermost; an operand whose
 * stride is 0 on either axis abstains; on conflicts C order wins.  Size-1
 * axes carry stride 0.  perm[0] is the innermost axis. */
static void axis_order(int ndim, const int64_t *shape, int nop, int64_t st[][ORA_MAXDIM], int *perm)
{
    int i0, i1, k, op;
    for (k = 0; k < ndim; k++) perm[k] = ndim - 1 - k;
    for (i0 = 1; i0 < ndim; i0++) {
        int ipos = i0, j0 = perm[i0];
        for (i1 = i0 - 1; i1 >= 0; i1--) {
            int ambig = 1, swap = 0, j1 = perm[i1];
            for (op = 0; op < nop; op++) {
                int64_t s0 = shape[j0] == 1 ? 0 : st[op][j0];
                int64_t s1 = shape[j1] == 1 ? 0 : st[op][j1];
                if (s0 != 0 && s1 != 0) {
                    if (llabs(s1) <= llabs(s0)) swap = 0;
                    else if (ambig) swap = 1;
                    ambig = 0;
                }
            }
            if (!ambig) {
                if (swap) ipos = i1;
                else break;
            }
        }
        if (ipos != i0) {
            for (k = i0; k > ipos; k--) perm[k] = perm[k - 1];
            perm[ipos] = j0;
        }
    }
}

/* K-order output strides of an elementwise call (lib.rs:1290
 * apply_ufunc_order 'K', array.rs:799): operands broadcast to `shape`,
 * output dense in iterator order.  If every operand is C-contiguous (or
 * every one F-contiguous) NumPy keeps that order outright. */
int ora_korder_strides(int nop, const ora_desc *const *ops, int ndim, const int64_t *shape, int64_t *out_strides)
{
    int64_t st[4][ORA_MAXDIM];
    int perm[ORA_MAXDIM], k, op;
    int64_t acc = 1;
    for (op = 0; op < nop; op++) bstrides(ops[op], ndim, shape, st[op]);
    axis_order(ndim, shape, nop, st, perm);
    for (k = 0; k < ndim; k++) {
        out_strides[perm[k]] = acc;
        acc *= shape[perm[k]] > 0 ? shape[perm[k]] : 1;
    }
    return ORA_OK;
}

static const char *elem_ptr(const ora_desc *d, const int64_t *idx, const int64_t *st, int n)
{
    int64_t off = 0;
    int i;
    for (i = 0; i < n; i++) off += idx[i] * st[i];
    return (const char *)d->data + off * ITEMSIZE[d->dtype];
}

static int64_t shape_size(int n, const int64_t *s)
{
    int64_t t = 1;
    int i;
    for (i = 0; i < n; i++) t *= s[i];
    return t;
}

/* binary_op (ufunc.rs:3628) -> binary_elementwise (ufunc.rs:3539): an
 * NdIter odometer over the broadcast output shape, casting each operand
 * element to the loop dtype and applying the scalar kernel. */
int ora_binary(int op, const ora_desc *a, const ora_desc *b, const ora_desc *out)
{
    int ldt, odt, rc, n = out->ndim, i;
    int64_t sa[ORA_MAXDIM], sb[ORA_MAXDIM], idx[ORA_MAXDIM], total;
    char ca[8], cb[8], co[8];
    rc = ora_binary_types(op, a->dtype, b->dtype, &ldt, &odt);
    if (rc) return rc;
    if (odt != out->dtype) return ORA_ETYPE;
    bstrides(a, n, out->shape, sa);
    bstrides(b, n, out->shape, sb);
    total = shape_size(n, out->shape);
    memset(idx, 0, sizeof idx);
    {
        int mixed_su = op >= OP_LESS && ldt == DT_F64 &&
                       ((a->dtype == DT_I64 && b->dtype == DT_U64) || (a->dtype == DT_U64 && b->dtype == DT_I64));
        int64_t e;
        for (e = 0; e < total; e++) {
            const char *pa = elem_ptr(a, idx, sa, n), *pb = elem_ptr(b, idx, sb, n);
            char *po = (char *)elem_ptr(out, idx, out->strides, n);
            if (mixed_su) {
                /* uint64 op int64 == int64 (swapped op) uint64 */
                int sw = op == OP_LESS ? OP_GREATER : op == OP_LESS_EQUAL ? OP_GREATER_EQUAL
                       : op == OP_GREATER ? OP_LESS : op == OP_GREATER_EQUAL ? OP_LESS_EQUAL : op;
                int r = a->dtype == DT_I64 ? cmp_su(op, *(const int64_t *)pa, *(const uint64_t *)pb)
                                           : cmp_su(sw, *(const int64_t *)pb, *(const uint64_t *)pa);
                *(uint8_t *)po = (uint8_t)r;
            } else {
                cast_elem(a->dtype, pa, ldt, ca);
                cast_elem(b->dtype, pb, ldt, cb);
                binary_elem(op, ldt, ca, cb, odt, co);
                memcpy(po, co, ITEMSIZE[odt]);
            }
            for (i = n - 1; i >= 0; i--) {
                if (++idx[i] < out->shape[i]) break;
                idx[i] = 0;
            }
        }
    }
    return ORA_OK;
}

/* ora_binary on every host core (the CPU baseline's multi-thread leg; the
 * reference itself is single-threaded): the output's first axis is cut
 * into one block of rows per thread and each block runs the same scalar
 * loop on sub-descriptors.  Elementwise results do not depend on the
 * split.  Returns the thread count used in *nthreads. */
int ora_binary_omp(int op, const ora_desc *a, const ora_desc *b, const ora_desc *out, int *nthreads)
{
    int n = out->ndim, ldt, odt, rc = ORA_OK, nt = 1;
    int64_t sa[ORA_MAXDIM], sb[ORA_MAXDIM], rows;
    rc = ora_binary_types(op, a->dtype, b->dtype, &ldt, &odt);
    if (rc) return rc;
    if (odt != out->dtype) return ORA_ETYPE;
    if (n == 0) {
        *nthreads = 1;
        return ora_binary(op, a, b, out);
    }
    bstrides(a, n, out->shape, sa);
    bstrides(b, n, out->shape, sb);
    rows = out->shape[0];
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    if (nt > rows) nt = rows > 0 ? (int)rows : 1;
    *nthreads = nt;
    {
        int t, bad = 0;
#pragma omp parallel for schedule(static) num_threads(nt) reduction(| : bad)
        for (t = 0; t < nt; t++) {
            const int64_t r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
            ora_desc da = *a, db = *b, dout = *out;
            da.ndim = db.ndim = n;
            memcpy(da.shape, out->shape, sizeof(int64_t) * n);
            memcpy(db.shape, out->shape, sizeof(int64_t) * n);
            memcpy(da.strides, sa, sizeof(int64_t) * n);
            memcpy(db.strides, sb, sizeof(int64_t) * n);
            da.shape[0] = db.shape[0] = dout.shape[0] = r1 - r0;
            da.data = (char *)a->data + r0 * sa[0] * ITEMSIZE[a->dtype];
            db.data = (char *)b->data + r0 * sb[0] * ITEMSIZE[b->dtype];
            dout.data = (char *)out->data + r0 * out->strides[0] * ITEMSIZE[out->dtype];
            if (r1 > r0) bad |= ora_binary(op, &da, &db, &dout);
        }
        if (bad) rc = ORA_EVALUE;
    }
    return rc;
}

/* where_select (sort.rs:1869-1914): out = cond ? x : y, three-way
 * broadcast, values cast to the output dtype. */
int ora_where(const ora_desc *c, const ora_desc *x, const ora_desc *y, const ora_desc *out)
{
    int n = out->ndim, i;
    int64_t sc[ORA_MAXDIM], sx[ORA_MAXDIM], sy[ORA_MAXDIM], idx[ORA_MAXDIM], total, e;
    char cv[8];
    bstrides(c, n, out->shape, sc);
    bstrides(x, n, out->shape, sx);
    bstrides(y, n, out->shape, sy);
    total = shape_size(n, out->shape);
    memset(idx, 0, sizeof idx);
    for (e = 0; e < total; e++) {
        const char *pc = elem_ptr(c, idx, sc, n);
        char *po = (char *)elem_ptr(out, idx, out->strides, n);
        cast_elem(c->dtype, pc, DT_BOOL, cv);
        if (cv[0]) cast_elem(x->dtype, elem_ptr(x, idx, sx, n), out->dtype, po);
        else cast_elem(y->dtype, elem_ptr(y, idx, sy, n), out->dtype, po);
        for (i = n - 1; i >= 0; i--) {
            if (++idx[i] < out->shape[i]) break;
            idx[i] = 0;
        }
    }
    return ORA_OK;
}

/* NdArray::cast_to (array.rs:393) -> Buffer::cast_to (buffer.rs:210) */
int ora_cast(const ora_desc *in, const ora_desc *out)
{
    int n = out->ndim, i;
    int64_t si[ORA_MAXDIM], idx[ORA_MAXDIM], total, e;
    bstrides(in, n, out->shape, si);
    total = shape_size(n, out->shape);
    memset(idx, 0, sizeof idx);
    for (e = 0; e < total; e++) {
        cast_elem(in->dtype, elem_ptr(in, idx, si, n), out->dtype, (char *)elem_ptr(out, idx, out->strides, n));
        for (i = n - 1; i >= 0; i--) {
            if (++idx[i] < out->shape[i]) break;
            idx[i] = 0;
        }
    }
    return ORA_OK;
}

/* ------------------------------------------------------------------------ */
/* IEEE exception flags (fpe_check_after_call, ionp-py/src/lib.rs:9077;     */
/* ionp-core/src/fpe.rs:173-360), as NumPy 2.2.6 reads them from the x86    */
/* status register: 1 divide by zero, 2 overflow, 4 underflow, 8 invalid.  */
/* ------------------------------------------------------------------------ */
static int fe_bits(void)
{
    int e = fetestexcept(FE_DIVBYZERO | FE_OVERFLOW | FE_UNDERFLOW | FE_INVALID), r = 0;
    if (e & FE_DIVBYZERO) r |= 1;
    if (e & FE_OVERFLOW) r |= 2;
    if (e & FE_UNDERFLOW) r |= 4;
    if (e & FE_INVALID) r |= 8;
    return r;
}

/* ora_binary under the host's own status register: the scalar loop's
 * arithmetic (plain C float ops on x86-64 SSE) raises exactly the flags
 * NumPy's loop raises.  Arithmetic ops only (add, subtract, multiply,
 * true_divide): the oracle's comparisons are ordered C compares, which
 * signal invalid on NaN where NumPy's do not. */
int ora_binary_fpe(int op, const ora_desc *a, const ora_desc *b, const ora_desc *out, int *flags)
{
    int rc;
    if (op > OP_TRUE_DIVIDE) return ORA_ENOLOOP;
    feclearexcept(FE_ALL_EXCEPT);
    rc = ora_binary(op, a, b, out);
    *flags = fe_bits();
    feclearexcept(FE_ALL_EXCEPT);
    return rc;
}

/* Flags of casting every element of `in` to dtype `to`: float -> integer
 * raises invalid when the conversion instruction the compiler uses does
 * not fit (cvttsd2si to 32 bits for int8..int32, to 64 bits for uint32 and
 * int64; for uint64 the two-step sequence of cvtt_u64); float64 -> float32
 * takes its overflow / underflow from the host's own conversion. */
int ora_cast_fpe(const ora_desc *in, int to, int *flags)
{
    int n = in->ndim, i, r = 0;
    int64_t idx[ORA_MAXDIM], total, e;
    if (!is_float_any(in->dtype) && to != DT_F16) {
        *flags = 0;
        return ORA_OK;
    }
    total = shape_size(n, in->shape);
    memset(idx, 0, sizeof idx);
    feclearexcept(FE_ALL_EXCEPT);
    for (e = 0; e < total; e++) {
        const char *p = elem_ptr(in, idx, in->strides, n);
        double x;
        if (to == DT_F16) { /* over / underflow raised by the conversion itself */
            char tmp[2];
            cast_elem(in->dtype, p, DT_F16, tmp);
            goto next;
        }
        x = in->dtype == DT_F32 ? (double)*(const float *)p
          : in->dtype == DT_F16 ? half_value(*(const uint16_t *)p) : *(const double *)p;
        switch (to) {
        case DT_I8: case DT_U8: case DT_I16: case DT_U16: case DT_I32:
            if (!(x > -2147483649.0 && x < 2147483648.0)) r |= 8;
            break;
        case DT_U32: case DT_I64:
            if (!(x >= -9223372036854775808.0 && x < 9223372036854775808.0)) r |= 8;
            break;
        case DT_U64:
            if (!(x >= -9223372036854775808.0 && x < 18446744073709551616.0)) r |= 8;
            break;
        case DT_F32:
            if (in->dtype == DT_F64) {
                volatile double vx = x;
                volatile float f = (float)vx;
                (void)f;
            }
            break;
        default:
            break;
        }
    next:
        for (i = n - 1; i >= 0; i--) {
            if (++idx[i] < in->shape[i]) break;
            idx[i] = 0;
        }
    }
    r |= fe_bits() & 6; /* over / under from the float64 -> float32 / float16 conversions */
    feclearexcept(FE_ALL_EXCEPT);
    *flags = r;
    return ORA_OK;
}

/* ------------------------------------------------------------------------ */
/* pairwise summation (ufunc.rs:6349-6448)                                  */
/* ------------------------------------------------------------------------ */
#define PW_BLOCKSIZE 128

/* Elements are fetched through a gather list of byte offsets so the same
 * tree serves strided and buffered windows. */
static double pw_f64(const char *base, const int64_t *off, int64_t n)
{
    if (n < 8) {
        double res = -0.0; /* NumPy 2.2.6's short leaf starts from -0.0, so a sum of -0.0 stays -0.0
                              (ufunc.rs:6424 starts from get(base): the same value) */
        int64_t i;
        for (i = 0; i < n; i++) res = f64_arith(OP_ADD, res, *(const double *)(base + off[i]));
        return res;
    } else if (n <= PW_BLOCKSIZE) {
        double r[8], res;
        int64_t i;
        int j;
        for (j = 0; j < 8; j++) r[j] = *(const double *)(base + off[j]);
        for (i = 8; i < n - (n % 8); i += 8)
            for (j = 0; j < 8; j++) r[j] = f64_arith(OP_ADD, r[j], *(const double *)(base + off[i + j]));
        res = f64_arith(OP_ADD, f64_arith(OP_ADD, f64_arith(OP_ADD, r[0], r[1]), f64_arith(OP_ADD, r[2], r[3])),
                        f64_arith(OP_ADD, f64_arith(OP_ADD, r[4], r[5]), f64_arith(OP_ADD, r[6], r[7])));
        for (; i < n; i++) res = f64_arith(OP_ADD, res, *(const double *)(base + off[i]));
        return res;
    } else {
        int64_t n2 = n / 2;
        n2 -= n2 % 8;
        return f64_arith(OP_ADD, pw_f64(base, off, n2), pw_f64(base, off + n2, n - n2));
    }
}
static float pw_f32(const char *base, const int64_t *off, int64_t n)
{
    if (n < 8) {
        float res = -0.0f;
        int64_t i;
        for (i = 0; i < n; i++) res = f32_arith(OP_ADD, res, *(const float *)(base + off[i]));
        return res;
    } else if (n <= PW_BLOCKSIZE) {
        float r[8], res;
        int64_t i;
        int j;
        for (j = 0; j < 8; j++) r[j] = *(const float *)(base + off[j]);
        for (i = 8; i < n - (n % 8); i += 8)
            for (j = 0; j < 8; j++) r[j] = f32_arith(OP_ADD, r[j], *(const float *)(base + off[i + j]));
        res = f32_arith(OP_ADD, f32_arith(OP_ADD, f32_arith(OP_ADD, r[0], r[1]), f32_arith(OP_ADD, r[2], r[3])),
                        f32_arith(OP_ADD, f32_arith(OP_ADD, r[4], r[5]), f32_arith(OP_ADD, r[6], r[7])));
        for (; i < n; i++) res = f32_arith(OP_ADD, res, *(const float *)(base + off[i]));
        return res;
    } else {
        int64_t n2 = n / 2;
        n2 -= n2 % 8;
        return f32_arith(OP_ADD, pw_f32(base, off, n2), pw_f32(base, off + n2, n - n2));
    }
}

/* ------------------------------------------------------------------------ */
/* maximum/minimum reduce loop models                                       */
/* ------------------------------------------------------------------------ */
/* In-order position of each SIMD lane in the AVX-512 horizontal reduce
 * (lanes listed first-to-last; ties go to the later lane).  Measured on
 * NumPy 2.2.6 (AVX512_SKX dispatch) with single -0 among +0 probes. */
static const int LANE_ORDER_F64[8] = {6, 2, 4, 0, 7, 3, 5, 1};
static const int LANE_ORDER_F32[16] = {12, 4, 8, 0, 13, 5, 9, 1, 14, 6, 10, 2, 15, 7, 11, 3};

/* Contiguous inner loop of maximum/minimum.reduce: acc lanes seeded with
 * the running value p, lane k folds elements k, k+V, ... with maxn; the
 * lanes are combined in LANE_ORDER (NaN anywhere -> canonical NaN); the
 * remaining n % V elements are folded in order with the scalar op. */
static double simd_ext_f64(int ismax, double p, const double *x, int64_t n)
{
    const int V = 8;
    double acc[8], r;
    int64_t nv = n - n % V, i;
    int k, anynan = 0;
    if (n < 1) return p;
    for (k = 0; k < V; k++) acc[k] = p;
    for (i = 0; i < nv; i += V)
        for (k = 0; k < V; k++) acc[k] = ismax ? f64_max(acc[k], x[i + k]) : f64_min(acc[k], x[i + k]);
    for (k = 0; k < V; k++) anynan |= acc[k] != acc[k];
    if (anynan) r = bits_f64(CANON64);
    else {
        r = acc[LANE_ORDER_F64[0]];
        for (k = 1; k < V; k++) r = ismax ? f64_max(r, acc[LANE_ORDER_F64[k]]) : f64_min(r, acc[LANE_ORDER_F64[k]]);
    }
    for (i = nv; i < n; i++) r = ismax ? f64_max(r, x[i]) : f64_min(r, x[i]);
    return r;
}
static float simd_ext_f32(int ismax, float p, const float *x, int64_t n)
{
    const int V = 16;
    float acc[16], r;
    int64_t nv = n - n % V, i;
    int k, anynan = 0;
    if (n < 1) return p;
    for (k = 0; k < V; k++) acc[k] = p;
    for (i = 0; i < nv; i += V)
        for (k = 0; k < V; k++) acc[k] = ismax ? f32_max(acc[k], x[i + k]) : f32_min(acc[k], x[i + k]);
    for (k = 0; k < V; k++) anynan |= acc[k] != acc[k];
    if (anynan) r = bits_f32(CANON32);
    else {
        r = acc[LANE_ORDER_F32[0]];
        for (k = 1; k < V; k++) r = ismax ? f32_max(r, acc[LANE_ORDER_F32[k]]) : f32_min(r, acc[LANE_ORDER_F32[k]]);
    }
    for (i = nv; i < n; i++) r = ismax ? f32_max(r, x[i]) : f32_min(r, x[i]);
    return r;
}

/* ------------------------------------------------------------------------ */
/* the reduction iteration model (ufunc.rs:7634 reduce_axis ->              */
/* 7091 reduce_axis_generic -> 6658 reduce_axis_pairwise_narrow_wide)       */
/* ------------------------------------------------------------------------ */
/*
 * Axes of size > 1 are ordered by the iterator rule (axis_order on the
 * input's strides).  The innermost run of reduced axes, W, is the "wide"
 * block; the remaining reduced axes are "outer".  For one output element:
 *
 *   acc = seed
 *   for each outer index (iterator order, innermost outer axis fastest):
 *       W's elements in iterator order are cut into windows of `bufsize`;
 *       acc = acc (+) inner(window)        [W empty: acc = acc (+) x]
 *
 * For add the seed is 0 and inner() is the pairwise tree.  For
 * maximum/minimum there is no identity: the first element of the group
 * seeds acc and is skipped, and inner() is the SIMD model on windows whose
 * elements are contiguous in memory (or gathered into NumPy's buffer),
 * otherwise an in-order scalar fold.
 */
typedef struct {
    int nw, nouter;
    int w_axes[ORA_MAXDIM];     /* innermost first */
    int outer_axes[ORA_MAXDIM]; /* innermost first */
    int64_t wsize;
    int64_t w_inner_len;        /* elements in W's innermost coalesced run */
    int64_t w_run_stride;       /* element stride of that run */
} red_plan;

/* mask: NULL, or the where= operand broadcast to in's shape; it takes part
 * in the iterator's axis order and in coalescing, as NumPy's third
 * reduction operand does */
static void make_red_plan_m(const ora_desc *in, const ora_desc *mask, const uint8_t *axmask, red_plan *p)
{
    int64_t st[2][ORA_MAXDIM];
    int perm[ORA_MAXDIM], k, n = in->ndim, inW = 1;
    memset(p, 0, sizeof *p);
    for (k = 0; k < n; k++) {
        st[0][k] = in->strides[k];
        st[1][k] = mask ? mask->strides[k] : 0;
    }
    axis_order(n, in->shape, mask ? 2 : 1, st, perm);
    p->wsize = 1;
    for (k = 0; k < n; k++) {
        int a = perm[k];
        if (in->shape[a] == 1) continue;
        if (axmask[a]) {
            if (inW) { p->w_axes[p->nw++] = a; p->wsize *= in->shape[a]; }
            else p->outer_axes[p->nouter++] = a;
        } else inW = 0;
    }
    /* W's innermost coalesced run (npyiter_coalesce_axes) */
    p->w_run_stride = p->nw ? in->strides[p->w_axes[0]] : 1;
    {
        int64_t run = p->nw ? in->shape[p->w_axes[0]] : 1;
        int64_t mrun_st = p->nw && mask ? mask->strides[p->w_axes[0]] : 0;
        for (k = 1; k < p->nw; k++) {
            int a = p->w_axes[k];
            if (in->strides[a] != p->w_run_stride * run) break;
            if (mask && mask->strides[a] != mrun_st * run) break;
            run *= in->shape[a];
        }
        p->w_inner_len = run;
    }
}
static void make_red_plan(const ora_desc *in, const uint8_t *axmask, red_plan *p)
{
    make_red_plan_m(in, NULL, axmask, p);
}

/* Byte offsets of W's elements (iterator order) relative to a base. */
static void w_offsets(const ora_desc *in, const red_plan *p, int64_t *off)
{
    int64_t idx[ORA_MAXDIM], e;
    int k, isz = ITEMSIZE[in->dtype];
    memset(idx, 0, sizeof idx);
    for (e = 0; e < p->wsize; e++) {
        int64_t o = 0;
        for (k = 0; k < p->nw; k++) o += idx[k] * in->strides[p->w_axes[k]];
        off[e] = o * isz;
        for (k = 0; k < p->nw; k++) {
            if (++idx[k] < in->shape[p->w_axes[k]]) break;
            idx[k] = 0;
        }
    }
}

/* one float16 step acc (op) v of the HALF loops (binary_elem's float16 loop) */
static uint16_t half_step(int op, uint16_t acc, uint16_t v)
{
    uint16_t r;
    binary_elem(op, DT_F16, (const char *)&acc, (const char *)&v, DT_F16, (char *)&r);
    return r;
}

/* reduce_axis with a float16 output, as NumPy's HALF loops evaluate it:
 * the reduce loop keeps a float32 accumulator for one call (one bufsize
 * window) and stores it back to the float16 output: add -- acc =
 * half(float(acc) + pw_f32(window)); multiply -- a float32 product chain
 * from float(acc), rounded once per window; maximum / minimum -- element by
 * element on the halves.  W empty: every element is its own call.  Inputs
 * of another dtype are cast to float16 first (NumPy's buffered cast). */
static int ora_reduce_half(int op, const ora_desc *in, const uint8_t *axmask, const ora_desc *out, int64_t bufsize,
                           const void *initial)
{
    red_plan p;
    int n = in->ndim, k, nk = 0, kept[ORA_MAXDIM], isz = ITEMSIZE[in->dtype];
    int64_t *woff, *foff, okidx[ORA_MAXDIM], nout = 1, o, gsize = 1, e;
    float *fbuf;
    uint16_t *hbuf;
    make_red_plan(in, axmask, &p);
    for (k = 0; k < n; k++) {
        if (!axmask[k]) kept[nk++] = k;
        else gsize *= in->shape[k];
    }
    for (k = 0; k < nk; k++) nout *= in->shape[kept[k]];
    if (gsize == 0 && (op == OP_MAXIMUM || op == OP_MINIMUM) && nout > 0 && !initial) return ORA_EVALUE;
    woff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(p.wsize > 0 ? p.wsize : 1));
    foff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(bufsize > 0 ? bufsize : 1));
    fbuf = (float *)malloc(sizeof(float) * (size_t)(bufsize > 0 ? bufsize : 1));
    hbuf = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)(bufsize > 0 ? bufsize : 1));
    for (e = 0; e < bufsize; e++) foff[e] = e * 4;
    if (p.nw) w_offsets(in, &p, woff);
    memset(okidx, 0, sizeof okidx);
    for (o = 0; o < nout; o++) {
        int64_t base = 0, ooff = 0, oidx[ORA_MAXDIM], noutr = 1, r;
        uint16_t acc = op == OP_MULTIPLY ? 0x3c00u : 0;
        int first = !initial && (op == OP_MAXIMUM || op == OP_MINIMUM);
        if (initial) memcpy(&acc, initial, 2);
        for (k = 0; k < nk; k++) {
            base += okidx[k] * in->strides[kept[k]];
            ooff += okidx[k] * out->strides[k];
        }
        for (k = 0; k < p.nouter; k++) noutr *= in->shape[p.outer_axes[k]];
        if (gsize == 0) noutr = 0;
        memset(oidx, 0, sizeof oidx);
        for (r = 0; r < noutr; r++) {
            int64_t ob = base, s0;
            const char *pb;
            for (k = 0; k < p.nouter; k++) ob += oidx[k] * in->strides[p.outer_axes[k]];
            pb = (const char *)in->data + ob * isz;
            for (s0 = 0; s0 < (p.nw ? p.wsize : 1); s0 += bufsize) {
                int64_t len = p.nw ? (p.wsize - s0 < bufsize ? p.wsize - s0 : bufsize) : 1;
                for (e = 0; e < len; e++) {
                    cast_elem(in->dtype, pb + (p.nw ? woff[s0 + e] : 0), DT_F16, (char *)&hbuf[e]);
                    fbuf[e] = half_f(hbuf[e]);
                }
                if (!p.nw || op == OP_MAXIMUM || op == OP_MINIMUM) {
                    for (e = 0; e < len; e++) {
                        acc = first ? hbuf[e] : half_step(op, acc, hbuf[e]);
                        first = 0;
                    }
                } else if (op == OP_ADD) {
                    acc = f32_to_half(f32_bits(f32_arith(OP_ADD, half_f(acc), pw_f32((const char *)fbuf, foff, len))));
                } else {
                    float f = half_f(acc);
                    for (e = 0; e < len; e++) f = f32_arith(OP_MULTIPLY, f, fbuf[e]);
                    acc = f32_to_half(f32_bits(f));
                }
            }
            for (k = 0; k < p.nouter; k++) {
                if (++oidx[k] < in->shape[p.outer_axes[k]]) break;
                oidx[k] = 0;
            }
        }
        memcpy((char *)out->data + ooff * 2, &acc, 2);
        for (k = nk - 1; k >= 0; k--) {
            if (++okidx[k] < in->shape[kept[k]]) break;
            okidx[k] = 0;
        }
    }
    free(woff);
    free(foff);
    free(fbuf);
    free(hbuf);
    return ORA_OK;
}

/* reduce op codes for ora_reduce: OP_ADD, OP_MAXIMUM, OP_MINIMUM */
int ora_reduce(int op, const ora_desc *in, const uint8_t *axmask, const ora_desc *out, int64_t bufsize,
               const void *initial)
{
    red_plan p;
    int n = in->ndim, k, nk = 0, kept[ORA_MAXDIM];
    int isz = ITEMSIZE[in->dtype], cdt;
    int64_t *woff, okidx[ORA_MAXDIM], nout, o, gsize = 1;
    char *buf = NULL;
    if (op != OP_ADD && op != OP_MULTIPLY && op != OP_MAXIMUM && op != OP_MINIMUM) return ORA_ENOLOOP;
    /* the loop runs in the output dtype (the default of ora_reduce_dtype or
       an explicit dtype=); elements are cast to it as they are read */
    cdt = out->dtype;
    if (cdt == DT_F16) return ora_reduce_half(op, in, axmask, out, bufsize, initial);
    /* a float loop reads float input as is: another float dtype is cast first
       (the caller's buffered cast); an integer / bool loop casts each element
       (reductions cast unsafely, as NumPy's ufunc.reduce does by default) */
    if (in->dtype == DT_F16 && is_float(cdt)) return ORA_EVALUE;
    if (is_float(cdt) && in->dtype != cdt) return ORA_EVALUE; /* float loops read the input as is: cast it first */
    make_red_plan(in, axmask, &p);
    for (k = 0; k < n; k++) {
        if (!axmask[k]) kept[nk++] = k;
        else gsize *= in->shape[k];
    }
    nout = 1;
    for (k = 0; k < nk; k++) nout *= in->shape[kept[k]];
    if (gsize == 0 && op != OP_ADD && op != OP_MULTIPLY && nout > 0 && !initial) return ORA_EVALUE; /* no identity */
    woff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(p.wsize > 0 ? p.wsize : 1));
    buf = (char *)malloc((size_t)(bufsize > 0 ? bufsize : 1) * 8);
    if (p.nw) w_offsets(in, &p, woff);
    memset(okidx, 0, sizeof okidx);
    for (o = 0; o < nout; o++) {
        int64_t base = 0, oidx[ORA_MAXDIM], noutr = 1, r;
        char *po;
        /* output element address: out has the kept axes in order */
        {
            int64_t ooff = 0;
            for (k = 0; k < nk; k++) ooff += okidx[k] * out->strides[k];
            po = (char *)out->data + ooff * ITEMSIZE[out->dtype];
        }
        for (k = 0; k < nk; k++) base += okidx[k] * in->strides[kept[k]];
        for (k = 0; k < p.nouter; k++) noutr *= in->shape[p.outer_axes[k]];
        memset(oidx, 0, sizeof oidx);
        if (gsize == 0) {
            if (initial) memcpy(po, initial, ITEMSIZE[cdt]);
            else if (op == OP_MULTIPLY) { /* multiply's identity 1 */
                char one[8];
                const uint8_t b1 = 1;
                cast_elem(DT_BOOL, (const char *)&b1, cdt, one);
                memcpy(po, one, ITEMSIZE[cdt]);
            } else memset(po, 0, ITEMSIZE[cdt]);
        } else if (!is_float(cdt)) {
            /* sequential_group (ufunc.rs:6546) on exact dtypes: order-free */
            uint64_t uacc = 0;
            int64_t sacc = 0, e;
            int first = 1, uns = cdt == DT_BOOL || is_unsigned_int(cdt);
            if (op == OP_MULTIPLY) { /* identity 1; wrapping products are order-free */
                uacc = 1;
                sacc = 1;
                first = 0;
            }
            if (initial) { /* initial= seeds the fold (reduce_axis, ufunc.rs:7634 `initial`) */
                uacc = (uint64_t)load_i(cdt, (const char *)initial);
                sacc = load_i(cdt, (const char *)initial);
                first = 0;
            }
            for (r = 0; r < noutr; r++) {
                int64_t ob = base;
                for (k = 0; k < p.nouter; k++) ob += oidx[k] * in->strides[p.outer_axes[k]];
                for (e = 0; e < p.wsize; e++) {
                    const char *pe = (const char *)in->data + ob * isz + (p.nw ? woff[e] : 0);
                    char tmp[8];
                    cast_elem(in->dtype, pe, cdt, tmp);
                    if (uns) {
                        uint64_t v = (uint64_t)load_i(cdt, tmp);
                        if (op == OP_ADD) uacc += v;
                        else if (op == OP_MULTIPLY) uacc *= v;
                        else if (first) uacc = v;
                        else if (op == OP_MAXIMUM) uacc = v > uacc ? v : uacc;
                        else uacc = v < uacc ? v : uacc;
                    } else {
                        int64_t v = load_i(cdt, tmp);
                        if (op == OP_ADD) sacc = (int64_t)((uint64_t)sacc + (uint64_t)v);
                        else if (op == OP_MULTIPLY) sacc = (int64_t)((uint64_t)sacc * (uint64_t)v);
                        else if (first) sacc = v;
                        else if (op == OP_MAXIMUM) sacc = v > sacc ? v : sacc;
                        else sacc = v < sacc ? v : sacc;
                    }
                    first = 0;
                }
                for (k = 0; k < p.nouter; k++) {
                    if (++oidx[k] < in->shape[p.outer_axes[k]]) break;
                    oidx[k] = 0;
                }
            }
            {
                uint64_t acc = uns ? uacc : (uint64_t)sacc;
                /* add on bool is logical or (bool_same, ufunc.rs:2341) */
                if (cdt == DT_BOOL) acc = acc != 0;
                switch (ITEMSIZE[cdt]) {
                case 1: *(uint8_t *)po = (uint8_t)acc; break;
                case 2: *(uint16_t *)po = (uint16_t)acc; break;
                case 4: *(uint32_t *)po = (uint32_t)acc; break;
                default: *(uint64_t *)po = acc; break;
                }
            }
        } else if (op == OP_MULTIPLY) {
            /* sequential_group_ordered (ufunc.rs:6583): no pairwise tree, one
               running product in iterator order (outer reduced index, then
               the W block), from initial (default 1) */
            double acc64 = initial && cdt == DT_F64 ? *(const double *)initial : 1.0;
            float acc32 = initial && cdt == DT_F32 ? *(const float *)initial : 1.0f;
            for (r = 0; r < noutr; r++) {
                int64_t ob = base, e;
                const char *pb;
                for (k = 0; k < p.nouter; k++) ob += oidx[k] * in->strides[p.outer_axes[k]];
                pb = (const char *)in->data + ob * isz;
                for (e = 0; e < p.wsize; e++) {
                    const char *pe = pb + (p.nw ? woff[e] : 0);
                    if (cdt == DT_F64) acc64 = f64_arith(OP_MULTIPLY, acc64, *(const double *)pe);
                    else acc32 = f32_arith(OP_MULTIPLY, acc32, *(const float *)pe);
                }
                for (k = 0; k < p.nouter; k++) {
                    if (++oidx[k] < in->shape[p.outer_axes[k]]) break;
                    oidx[k] = 0;
                }
            }
            if (cdt == DT_F64) *(double *)po = acc64;
            else *(float *)po = acc32;
        } else if (op == OP_ADD) {
            /* acc = initial (default 0); acc = acc + pw(window) per window */
            double acc64 = initial && cdt == DT_F64 ? *(const double *)initial : 0.0;
            float acc32 = initial && cdt == DT_F32 ? *(const float *)initial : 0.0f;
            for (r = 0; r < noutr; r++) {
                int64_t ob = base, s;
                const char *pb;
                for (k = 0; k < p.nouter; k++) ob += oidx[k] * in->strides[p.outer_axes[k]];
                pb = (const char *)in->data + ob * isz;
                if (!p.nw) {
                    if (cdt == DT_F64) acc64 = f64_arith(OP_ADD, acc64, *(const double *)pb);
                    else acc32 = f32_arith(OP_ADD, acc32, *(const float *)pb);
                } else {
                    for (s = 0; s < p.wsize; s += bufsize) {
                        int64_t len = p.wsize - s < bufsize ? p.wsize - s : bufsize;
                        if (cdt == DT_F64) acc64 = f64_arith(OP_ADD, acc64, pw_f64(pb, woff + s, len));
                        else acc32 = f32_arith(OP_ADD, acc32, pw_f32(pb, woff + s, len));
                    }
                }
                for (k = 0; k < p.nouter; k++) {
                    if (++oidx[k] < in->shape[p.outer_axes[k]]) break;
                    oidx[k] = 0;
                }
            }
            if (cdt == DT_F64) *(double *)po = acc64;
            else *(float *)po = acc32;
        } else {
            /* maximum / minimum on floats */
            /* without initial= the first element seeds the fold and is not
               fed to the loop; with it every element is */
            int ismax = op == OP_MAXIMUM, first = !initial;
            double acc64 = initial && cdt == DT_F64 ? *(const double *)initial : 0.0;
            float acc32 = initial && cdt == DT_F32 ? *(const float *)initial : 0.0f;
            for (r = 0; r < noutr; r++) {
                int64_t ob = base, s;
                const char *pb;
                for (k = 0; k < p.nouter; k++) ob += oidx[k] * in->strides[p.outer_axes[k]];
                pb = (const char *)in->data + ob * isz;
                if (!p.nw) {
                    if (cdt == DT_F64) {
                        double v = *(const double *)pb;
                        acc64 = first ? v : (ismax ? f64_max(acc64, v) : f64_min(acc64, v));
                    } else {
                        float v = *(const float *)pb;
                        acc32 = first ? v : (ismax ? f32_max(acc32, v) : f32_min(acc32, v));
                    }
                    first = 0;
                } else {
                    for (s = 0; s < p.wsize; s += bufsize) {
                        int64_t len = p.wsize - s < bufsize ? p.wsize - s : bufsize, e0 = s, e;
                        int contig;
                        /* A window inside W's innermost coalesced run is
                         * read in place with that run's stride; a window
                         * spanning runs is gathered into NumPy's contiguous
                         * buffer.  Decided on the whole window, before the
                         * first-visit skip below. */
                        if ((s / p.w_inner_len) == ((s + len - 1) / p.w_inner_len))
                            contig = p.w_run_stride == 1;
                        else
                            contig = 1;
                        if (first) {
                            if (cdt == DT_F64) acc64 = *(const double *)(pb + woff[0]);
                            else acc32 = *(const float *)(pb + woff[0]);
                            e0 = s + 1; len -= 1; first = 0;
                        }
                        if (len <= 0) continue;
                        if (contig) {
                            for (e = 0; e < len; e++)
                                memcpy(buf + e * isz, pb + woff[e0 + e], isz);
                            if (cdt == DT_F64) acc64 = simd_ext_f64(ismax, acc64, (const double *)buf, len);
                            else acc32 = simd_ext_f32(ismax, acc32, (const float *)buf, len);
                        } else {
                            /* strided loop: 8 scalar accumulators over
                             * blocks of 8, combined in accumulator order,
                             * folded into acc, then the len % 8 tail */
                            int64_t nb = len >= 8 ? len - len % 8 : 0;
                            if (cdt == DT_F64) {
                                if (nb) {
                                    double m[8], t;
                                    int j;
                                    for (j = 0; j < 8; j++) m[j] = *(const double *)(pb + woff[e0 + j]);
                                    for (e = 8; e < nb; e += 8)
                                        for (j = 0; j < 8; j++) {
                                            double v = *(const double *)(pb + woff[e0 + e + j]);
                                            m[j] = ismax ? f64_max(m[j], v) : f64_min(m[j], v);
                                        }
                                    t = m[0];
                                    for (j = 1; j < 8; j++) t = ismax ? f64_max(t, m[j]) : f64_min(t, m[j]);
                                    acc64 = ismax ? f64_max(acc64, t) : f64_min(acc64, t);
                                }
                                for (e = nb; e < len; e++) {
                                    double v = *(const double *)(pb + woff[e0 + e]);
                                    acc64 = ismax ? f64_max(acc64, v) : f64_min(acc64, v);
                                }
                            } else {
                                if (nb) {
                                    float m[8], t;
                                    int j;
                                    for (j = 0; j < 8; j++) m[j] = *(const float *)(pb + woff[e0 + j]);
                                    for (e = 8; e < nb; e += 8)
                                        for (j = 0; j < 8; j++) {
                                            float v = *(const float *)(pb + woff[e0 + e + j]);
                                            m[j] = ismax ? f32_max(m[j], v) : f32_min(m[j], v);
                                        }
                                    t = m[0];
                                    for (j = 1; j < 8; j++) t = ismax ? f32_max(t, m[j]) : f32_min(t, m[j]);
                                    acc32 = ismax ? f32_max(acc32, t) : f32_min(acc32, t);
                                }
                                for (e = nb; e < len; e++) {
                                    float v = *(const float *)(pb + woff[e0 + e]);
                                    acc32 = ismax ? f32_max(acc32, v) : f32_min(acc32, v);
                                }
                            }
                        }
                    }
                }
                for (k = 0; k < p.nouter; k++) {
                    if (++oidx[k] < in->shape[p.outer_axes[k]]) break;
                    oidx[k] = 0;
                }
            }
            if (cdt == DT_F64) *(double *)po = acc64;
            else *(float *)po = acc32;
        }
        for (k = nk - 1; k >= 0; k--) {
            if (++okidx[k] < in->shape[kept[k]]) break;
            okidx[k] = 0;
        }
    }
    free(woff);
    free(buf);
    return ORA_OK;
}

/* Byte offsets of W's elements (iterator order) in the mask operand. */
static void w_mask_offsets(const ora_desc *m, const red_plan *p, int64_t *off)
{
    int64_t idx[ORA_MAXDIM], e;
    int k;
    memset(idx, 0, sizeof idx);
    for (e = 0; e < p->wsize; e++) {
        int64_t o = 0;
        for (k = 0; k < p->nw; k++) o += idx[k] * m->strides[p->w_axes[k]];
        off[e] = o;
        for (k = 0; k < p->nw; k++) {
            if (++idx[k] < m->shape[p->w_axes[k]]) break;
            idx[k] = 0;
        }
    }
}

/* Strided (non-SIMD) maximum/minimum inner loop on n elements at byte
 * offsets off[]: 8 scalar accumulators over blocks of 8 combined in order
 * and folded into p, then the n % 8 tail in order (the model ora_reduce
 * uses for windows read in place with a stride other than 1). */
static double strided_ext_f64(int ismax, double p, const char *pb, const int64_t *off, int64_t n)
{
    int64_t nb = n >= 8 ? n - n % 8 : 0, e;
    if (nb) {
        double m[8], t;
        int j;
        for (j = 0; j < 8; j++) m[j] = *(const double *)(pb + off[j]);
        for (e = 8; e < nb; e += 8)
            for (j = 0; j < 8; j++) {
                double v = *(const double *)(pb + off[e + j]);
                m[j] = ismax ? f64_max(m[j], v) : f64_min(m[j], v);
            }
        t = m[0];
        for (j = 1; j < 8; j++) t = ismax ? f64_max(t, m[j]) : f64_min(t, m[j]);
        p = ismax ? f64_max(p, t) : f64_min(p, t);
    }
    for (e = nb; e < n; e++) {
        double v = *(const double *)(pb + off[e]);
        p = ismax ? f64_max(p, v) : f64_min(p, v);
    }
    return p;
}
static float strided_ext_f32(int ismax, float p, const char *pb, const int64_t *off, int64_t n)
{
    int64_t nb = n >= 8 ? n - n % 8 : 0, e;
    if (nb) {
        float m[8], t;
        int j;
        for (j = 0; j < 8; j++) m[j] = *(const float *)(pb + off[j]);
        for (e = 8; e < nb; e += 8)
            for (j = 0; j < 8; j++) {
                float v = *(const float *)(pb + off[e + j]);
                m[j] = ismax ? f32_max(m[j], v) : f32_min(m[j], v);
            }
        t = m[0];
        for (j = 1; j < 8; j++) t = ismax ? f32_max(t, m[j]) : f32_min(t, m[j]);
        p = ismax ? f32_max(p, t) : f32_min(p, t);
    }
    for (e = nb; e < n; e++) {
        float v = *(const float *)(pb + off[e]);
        p = ismax ? f32_max(p, v) : f32_min(p, v);
    }
    return p;
}

/* reduce_axis with a where= mask (ufunc.rs:7172-7540 reduce_axis_masked),
 * as NumPy 2.2.6 evaluates it: the fold starts from initial= (else the
 * op's identity: add 0, multiply 1; maximum/minimum without initial= are
 * an error whatever the sizes), the mask takes part in the iterator's axis
 * order, and the reduce inner loop is called once per maximal run of True
 * mask values inside each `bufsize` window (NumPy's masked-loop wrapper):
 *   float add       acc = acc + pw(run)          (per run, not per window)
 *   float multiply  acc = acc * x, every selected x in order
 *   float max/min   the window's SIMD (or strided) loop model on the run
 *   integers        order-free fold of the selected elements
 * W empty: acc = op(acc, x) for each selected x in iterator order.
 * float16 output: the HALF loop per call -- add acc = half(float(acc) +
 * pw_f32(run)), multiply a float32 chain from float(acc) rounded at the
 * run's end, maximum / minimum element by element on the halves
 * (reduce_axis_masked_f16_widen, ufunc.rs:7414, widens the same way).
 * `mask` is a bool operand broadcast to in's shape (zero strides allowed).
 * The reference scans runs over whole W blocks (no bufsize windows); the
 * oracle follows NumPy, as for the unmasked fold. */
int ora_reduce_where(int op, const ora_desc *in, const uint8_t *axmask, const ora_desc *mask, const ora_desc *out,
                     int64_t bufsize, const void *initial)
{
    red_plan p;
    int n = in->ndim, k, nk = 0, kept[ORA_MAXDIM];
    int isz = ITEMSIZE[in->dtype], cdt = out->dtype, ismax = op == OP_MAXIMUM;
    int64_t *woff, *moff, *foff, okidx[ORA_MAXDIM], nout, o;
    char *buf;
    if (op != OP_ADD && op != OP_MULTIPLY && op != OP_MAXIMUM && op != OP_MINIMUM) return ORA_ENOLOOP;
    if (mask->dtype != DT_BOOL || mask->ndim != n) return ORA_ETYPE;
    if (in->dtype == DT_F16 && is_float(cdt)) return ORA_EVALUE; /* a float32 / float64 loop: cast it first */
    if (is_float(cdt) && in->dtype != cdt) return ORA_EVALUE; /* float loops read the input as is */
    if ((op == OP_MAXIMUM || op == OP_MINIMUM) && !initial) return ORA_EVALUE; /* no identity */
    make_red_plan_m(in, mask, axmask, &p);
    for (k = 0; k < n; k++)
        if (!axmask[k]) kept[nk++] = k;
    nout = 1;
    for (k = 0; k < nk; k++) nout *= in->shape[kept[k]];
    if (in->ndim && shape_size(n, in->shape) == 0) p.wsize = 0;
    woff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(p.wsize > 0 ? p.wsize : 1));
    moff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(p.wsize > 0 ? p.wsize : 1));
    buf = (char *)malloc((size_t)(bufsize > 0 ? bufsize : 1) * 8);
    foff = (int64_t *)malloc(sizeof(int64_t) * (size_t)(bufsize > 0 ? bufsize : 1));
    for (o = 0; o < bufsize; o++) foff[o] = o * 4;
    if (p.nw && p.wsize) {
        w_offsets(in, &p, woff);
        w_mask_offsets(mask, &p, moff);
    }
    memset(okidx, 0, sizeof okidx);
    for (o = 0; o < nout; o++) {
        int64_t base = 0, mbase = 0, oidx[ORA_MAXDIM], noutr = 1, r, ooff = 0;
        char *po;
        double a64;
        float a32;
        uint64_t ua;
        int64_t sa;
        uint16_t ah = op == OP_MULTIPLY ? 0x3c00u : 0; /* float16 output: the accumulator's bits */
        int uns = cdt == DT_BOOL || is_unsigned_int(cdt);
        for (k = 0; k < nk; k++) ooff += okidx[k] * out->strides[k];
        po = (char *)out->data + ooff * ITEMSIZE[cdt];
        for (k = 0; k < nk; k++) {
            base += okidx[k] * in->strides[kept[k]];
            mbase += okidx[k] * mask->strides[kept[k]];
        }
        for (k = 0; k < p.nouter; k++) noutr *= in->shape[p.outer_axes[k]];
        if (shape_size(n, in->shape) == 0) noutr = 0;
        /* seed: initial=, else the identity */
        a64 = op == OP_MULTIPLY ? 1.0 : 0.0;
        a32 = op == OP_MULTIPLY ? 1.0f : 0.0f;
        ua = op == OP_MULTIPLY ? 1 : 0;
        sa = op == OP_MULTIPLY ? 1 : 0;
        if (initial) {
            if (cdt == DT_F64) a64 = *(const double *)initial;
            else if (cdt == DT_F32) a32 = *(const float *)initial;
            else if (cdt == DT_F16) memcpy(&ah, initial, 2);
            else {
                ua = (uint64_t)load_i(cdt, (const char *)initial);
                sa = load_i(cdt, (const char *)initial);
            }
        }
        memset(oidx, 0, sizeof oidx);
        for (r = 0; r < noutr; r++) {
            int64_t ob = base, mb = mbase, s, e;
            const char *pb;
            const uint8_t *mp;
            for (k = 0; k < p.nouter; k++) {
                ob += oidx[k] * in->strides[p.outer_axes[k]];
                mb += oidx[k] * mask->strides[p.outer_axes[k]];
            }
            pb = (const char *)in->data + ob * isz;
            mp = (const uint8_t *)mask->data + mb;
            for (s = 0; s < (p.nw ? p.wsize : 1); s += bufsize) {
                int64_t len = p.nw ? (p.wsize - s < bufsize ? p.wsize - s : bufsize) : 1, a = 0, b;
                int contig;
                if (!p.nw) contig = 1;
                else if ((s / p.w_inner_len) == ((s + len - 1) / p.w_inner_len)) contig = p.w_run_stride == 1;
                else contig = 1;
                while (a < len) {
                    const int64_t *wo = p.nw ? woff + s + a : NULL;
                    static const int64_t zero = 0;
                    if (!mp[p.nw ? moff[s + a] : 0]) { a++; continue; }
                    b = a + 1;
                    while (b < len && mp[moff[s + b]]) b++;
                    if (!wo) wo = &zero;
                    /* one call of the inner loop on elements [a, b) */
                    if (cdt == DT_F16) {
                        /* NumPy's HALF loop: a float32 accumulator for the
                           call, stored back to float16 at its end (add:
                           float(acc) + pw_f32(run); multiply: a float32
                           chain); maximum / minimum element by element on
                           the halves; W empty: every element its own call */
                        float *fv = (float *)buf;
                        uint16_t h;
                        if (op == OP_ADD && p.nw) {
                            for (e = 0; e < b - a; e++) {
                                cast_elem(in->dtype, pb + wo[e], DT_F16, (char *)&h);
                                fv[e] = half_f(h);
                            }
                            ah = f32_to_half(f32_bits(f32_arith(OP_ADD, half_f(ah), pw_f32(buf, foff, b - a))));
                        } else if (op == OP_MULTIPLY && p.nw) {
                            float f = half_f(ah);
                            for (e = 0; e < b - a; e++) {
                                cast_elem(in->dtype, pb + wo[e], DT_F16, (char *)&h);
                                f = f32_arith(OP_MULTIPLY, f, half_f(h));
                            }
                            ah = f32_to_half(f32_bits(f));
                        } else {
                            for (e = 0; e < b - a; e++) {
                                cast_elem(in->dtype, pb + wo[e], DT_F16, (char *)&h);
                                ah = half_step(op, ah, h);
                            }
                        }
                    } else if (cdt == DT_F64 || cdt == DT_F32) {
                        if (op == OP_ADD && p.nw) {
                            if (cdt == DT_F64) a64 = f64_arith(OP_ADD, a64, pw_f64(pb, wo, b - a));
                            else a32 = f32_arith(OP_ADD, a32, pw_f32(pb, wo, b - a));
                        } else if (op == OP_ADD || op == OP_MULTIPLY) {
                            for (e = 0; e < b - a; e++) {
                                if (cdt == DT_F64) a64 = f64_arith(op, a64, *(const double *)(pb + wo[e]));
                                else a32 = f32_arith(op, a32, *(const float *)(pb + wo[e]));
                            }
                        } else if (!p.nw) { /* W empty: the elementwise loop, scalar max/min */
                            if (cdt == DT_F64) {
                                double v = *(const double *)pb;
                                a64 = ismax ? f64_max(a64, v) : f64_min(a64, v);
                            } else {
                                float v = *(const float *)pb;
                                a32 = ismax ? f32_max(a32, v) : f32_min(a32, v);
                            }
                        } else if (contig) {
                            for (e = 0; e < b - a; e++) memcpy(buf + e * isz, pb + wo[e], isz);
                            if (cdt == DT_F64) a64 = simd_ext_f64(ismax, a64, (const double *)buf, b - a);
                            else a32 = simd_ext_f32(ismax, a32, (const float *)buf, b - a);
                        } else {
                            if (cdt == DT_F64) a64 = strided_ext_f64(ismax, a64, pb, wo, b - a);
                            else a32 = strided_ext_f32(ismax, a32, pb, wo, b - a);
                        }
                    } else {
                        for (e = 0; e < b - a; e++) {
                            char tmp[8];
                            cast_elem(in->dtype, pb + wo[e], cdt, tmp);
                            if (uns) {
                                uint64_t v = (uint64_t)load_i(cdt, tmp);
                                ua = op == OP_ADD ? ua + v : op == OP_MULTIPLY ? ua * v
                                   : op == OP_MAXIMUM ? (v > ua ? v : ua) : (v < ua ? v : ua);
                            } else {
                                int64_t v = load_i(cdt, tmp);
                                sa = op == OP_ADD ? (int64_t)((uint64_t)sa + (uint64_t)v)
                                   : op == OP_MULTIPLY ? (int64_t)((uint64_t)sa * (uint64_t)v)
                                   : op == OP_MAXIMUM ? (v > sa ? v : sa) : (v < sa ? v : sa);
                            }
                        }
                    }
                    a = b;
                }
            }
            for (k = 0; k < p.nouter; k++) {
                if (++oidx[k] < in->shape[p.outer_axes[k]]) break;
                oidx[k] = 0;
            }
        }
        if (cdt == DT_F64) *(double *)po = a64;
        else if (cdt == DT_F32) *(float *)po = a32;
        else if (cdt == DT_F16) memcpy(po, &ah, 2);
        else {
            uint64_t acc = uns ? ua : (uint64_t)sa;
            if (cdt == DT_BOOL) acc = acc != 0; /* add: or; multiply: and */
            switch (ITEMSIZE[cdt]) {
            case 1: *(uint8_t *)po = (uint8_t)acc; break;
            case 2: *(uint16_t *)po = (uint16_t)acc; break;
            case 4: *(uint32_t *)po = (uint32_t)acc; break;
            default: *(uint64_t *)po = acc; break;
            }
        }
        for (k = nk - 1; k >= 0; k--) {
            if (++okidx[k] < in->shape[kept[k]]) break;
            okidx[k] = 0;
        }
    }
    free(woff);
    free(moff);
    free(foff);
    free(buf);
    return ORA_OK;
}

/* ufunc.reduceat (reduceat_binary_axis, ufunc.rs:8298-8460) as NumPy 2.2.6
 * evaluates it: along `axis`, output k of every fiber is x[i_k] when
 * i_{k+1} <= i_k (the last index runs to the end), else the reduce loop
 * seeded with x[i_k] on the rest of the segment [i_k + 1, i_{k+1}) as one
 * call (no buffer windows):
 *   float add       x[i_k] + pw(rest)   (the reference folds sequentially)
 *   float multiply  sequential product
 *   float max/min   the SIMD loop model on a contiguous rest (stride 1),
 *                   the strided loop model otherwise
 *   integers        order-free fold (in the output dtype)
 * `out` has the input's shape with axis `axis` of length nidx (any
 * strides); indices are checked by the caller (0 <= i < n). */
int ora_reduceat(int op, const ora_desc *in, int axis, const int64_t *idx, int64_t nidx, const ora_desc *out)
{
    int n = in->ndim, k, cdt = out->dtype, isz = ITEMSIZE[in->dtype], ismax = op == OP_MAXIMUM;
    int64_t len, st, ost, nfib = 1, f, i, fidx[ORA_MAXDIM];
    int64_t *off;
    char *buf;
    if (op != OP_ADD && op != OP_MULTIPLY && op != OP_MAXIMUM && op != OP_MINIMUM) return ORA_ENOLOOP;
    if (axis < 0 || axis >= n || out->ndim != n) return ORA_EAXIS;
    if (is_float(cdt) && in->dtype != cdt) return ORA_EVALUE;
    len = in->shape[axis];
    st = in->strides[axis];
    ost = out->strides[axis];
    for (k = 0; k < n; k++)
        if (k != axis) nfib *= in->shape[k];
    for (i = 0; i < nidx; i++)
        if (idx[i] < 0 || idx[i] >= len) return ORA_EVALUE;
    if (nidx == 0 || nfib == 0) return ORA_OK;
    off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(len > 0 ? len : 1));
    buf = (char *)malloc((size_t)(len > 0 ? len : 1) * 8);
    for (i = 0; i < len; i++) off[i] = i * st * isz;
    memset(fidx, 0, sizeof fidx);
    for (f = 0; f < nfib; f++) {
        int64_t ib = 0, ob = 0, q;
        for (k = 0; k < n; k++)
            if (k != axis) {
                ib += fidx[k] * in->strides[k];
                ob += fidx[k] * out->strides[k];
            }
        for (q = 0; q < nidx; q++) {
            const int64_t a = idx[q], b = q + 1 < nidx ? idx[q + 1] : len;
            const char *pb = (const char *)in->data + ib * isz;
            char *po = (char *)out->data + (ob + q * ost) * ITEMSIZE[cdt];
            const int64_t m = b > a ? b - a - 1 : 0; /* elements after the seed */
            if (cdt == DT_F16) { /* the HALF loop: one call seeded with x[a], float32 inside */
                uint16_t acc, v;
                int64_t e;
                cast_elem(in->dtype, pb + off[a], DT_F16, (char *)&acc);
                if (m > 0 && op == OP_ADD) {
                    float *fb = (float *)buf;
                    int64_t *fo = (int64_t *)malloc(sizeof(int64_t) * (size_t)m);
                    for (e = 0; e < m; e++) {
                        cast_elem(in->dtype, pb + off[a + 1 + e], DT_F16, (char *)&v);
                        fb[e] = half_f(v);
                        fo[e] = e * 4;
                    }
                    acc = f32_to_half(f32_bits(f32_arith(OP_ADD, half_f(acc), pw_f32(buf, fo, m))));
                    free(fo);
                } else if (m > 0 && op == OP_MULTIPLY) {
                    float p = half_f(acc);
                    for (e = 0; e < m; e++) {
                        cast_elem(in->dtype, pb + off[a + 1 + e], DT_F16, (char *)&v);
                        p = f32_arith(OP_MULTIPLY, p, half_f(v));
                    }
                    acc = f32_to_half(f32_bits(p));
                } else {
                    for (e = 0; e < m; e++) {
                        cast_elem(in->dtype, pb + off[a + 1 + e], DT_F16, (char *)&v);
                        acc = half_step(op, acc, v);
                    }
                }
                memcpy(po, &acc, 2);
            } else if (cdt == DT_F64 || cdt == DT_F32) {
                double a64 = cdt == DT_F64 ? *(const double *)(pb + off[a]) : 0.0;
                float a32 = cdt == DT_F32 ? *(const float *)(pb + off[a]) : 0.0f;
                if (m > 0) {
                    const int64_t *ro = off + a + 1;
                    int64_t e;
                    if (op == OP_ADD) {
                        if (cdt == DT_F64) a64 = f64_arith(OP_ADD, a64, pw_f64(pb, ro, m));
                        else a32 = f32_arith(OP_ADD, a32, pw_f32(pb, ro, m));
                    } else if (op == OP_MULTIPLY) {
                        for (e = 0; e < m; e++) {
                            if (cdt == DT_F64) a64 = f64_arith(OP_MULTIPLY, a64, *(const double *)(pb + ro[e]));
                            else a32 = f32_arith(OP_MULTIPLY, a32, *(const float *)(pb + ro[e]));
                        }
                    } else if (st == 1) {
                        for (e = 0; e < m; e++) memcpy(buf + e * isz, pb + ro[e], isz);
                        if (cdt == DT_F64) a64 = simd_ext_f64(ismax, a64, (const double *)buf, m);
                        else a32 = simd_ext_f32(ismax, a32, (const float *)buf, m);
                    } else {
                        if (cdt == DT_F64) a64 = strided_ext_f64(ismax, a64, pb, ro, m);
                        else a32 = strided_ext_f32(ismax, a32, pb, ro, m);
                    }
                }
                if (cdt == DT_F64) *(double *)po = a64;
                else *(float *)po = a32;
            } else {
                int uns = cdt == DT_BOOL || is_unsigned_int(cdt);
                char tmp[8];
                uint64_t ua;
                int64_t sa, e;
                cast_elem(in->dtype, pb + off[a], cdt, tmp);
                ua = (uint64_t)load_i(cdt, tmp);
                sa = load_i(cdt, tmp);
                for (e = 0; e < m; e++) {
                    cast_elem(in->dtype, pb + off[a + 1 + e], cdt, tmp);
                    if (uns) {
                        uint64_t v = (uint64_t)load_i(cdt, tmp);
                        ua = op == OP_ADD ? ua + v : op == OP_MULTIPLY ? ua * v
                           : op == OP_MAXIMUM ? (v > ua ? v : ua) : (v < ua ? v : ua);
                    } else {
                        int64_t v = load_i(cdt, tmp);
                        sa = op == OP_ADD ? (int64_t)((uint64_t)sa + (uint64_t)v)
                           : op == OP_MULTIPLY ? (int64_t)((uint64_t)sa * (uint64_t)v)
                           : op == OP_MAXIMUM ? (v > sa ? v : sa) : (v < sa ? v : sa);
                    }
                }
                {
                    uint64_t acc = uns ? ua : (uint64_t)sa;
                    if (cdt == DT_BOOL) acc = acc != 0; /* add: or; multiply: and */
                    switch (ITEMSIZE[cdt]) {
                    case 1: *(uint8_t *)po = (uint8_t)acc; break;
                    case 2: *(uint16_t *)po = (uint16_t)acc; break;
                    case 4: *(uint32_t *)po = (uint32_t)acc; break;
                    default: *(uint64_t *)po = acc; break;
                    }
                }
            }
        }
        for (k = n - 1; k >= 0; k--) {
            if (k == axis) continue;
            if (++fidx[k] < in->shape[k]) break;
            fidx[k] = 0;
        }
    }
    free(off);
    free(buf);
    return ORA_OK;
}

/* accumulate_axis (ufunc.rs:7967) -> accumulate_axis_generic: along `axis`,
 * out[0] = x[0] and out[i] = op(out[i-1], x[i]), in the output dtype (the
 * reduce-family compute dtype, or an explicit dtype=); elements are cast to
 * it as they are read.  Float add is the plain sequential sum (no pairwise
 * tree); maximum/minimum use the scalar rules of f64_max/f64_min; integers
 * wrap; on bool, add/maximum are logical or and minimum is logical and. */
int ora_accumulate(int op, const ora_desc *in, int axis, const ora_desc *out)
{
    int n = in->ndim, k, cdt = out->dtype, osz;
    int64_t idx[ORA_MAXDIM], len, nfib = 1, f, i;
    if (op != OP_ADD && op != OP_MULTIPLY && op != OP_MAXIMUM && op != OP_MINIMUM) return ORA_ENOLOOP;
    if (axis < 0 || axis >= (n > 0 ? n : 1)) return ORA_EAXIS;
    if (out->ndim != n) return ORA_EVALUE;
    for (k = 0; k < n; k++)
        if (out->shape[k] != in->shape[k]) return ORA_EVALUE;
    len = n > 0 ? in->shape[axis] : 1;
    for (k = 0; k < n; k++)
        if (k != axis) nfib *= in->shape[k];
    osz = ITEMSIZE[cdt];
    memset(idx, 0, sizeof idx);
    for (f = 0; f < nfib; f++) {
        int64_t ib = 0, ob = 0;
        char acc[8], v[8];
        for (k = 0; k < n; k++) {
            ib += idx[k] * in->strides[k];
            ob += idx[k] * out->strides[k];
        }
        for (i = 0; i < len; i++) {
            const char *pi = (const char *)in->data + (ib + i * (n > 0 ? in->strides[axis] : 0)) * ITEMSIZE[in->dtype];
            char *po = (char *)out->data + (ob + i * (n > 0 ? out->strides[axis] : 0)) * osz;
            cast_elem(in->dtype, pi, cdt, v);
            if (i == 0) {
                memcpy(acc, v, osz);
            } else if (cdt == DT_F16) {
                uint16_t a, b;
                memcpy(&a, acc, 2);
                memcpy(&b, v, 2);
                a = half_step(op, a, b);
                memcpy(acc, &a, 2);
            } else if (cdt == DT_F64) {
                double a = *(double *)acc, b = *(double *)v;
                a = op == OP_ADD || op == OP_MULTIPLY ? f64_arith(op, a, b) : op == OP_MAXIMUM ? f64_max(a, b) : f64_min(a, b);
                memcpy(acc, &a, 8);
            } else if (cdt == DT_F32) {
                float a = *(float *)acc, b = *(float *)v;
                a = op == OP_ADD || op == OP_MULTIPLY ? f32_arith(op, a, b) : op == OP_MAXIMUM ? f32_max(a, b) : f32_min(a, b);
                memcpy(acc, &a, 4);
            } else if (cdt == DT_BOOL) {
                uint8_t a = *(uint8_t *)acc != 0, b = *(uint8_t *)v != 0;
                a = op == OP_MINIMUM || op == OP_MULTIPLY ? (a & b) : (a | b);
                memcpy(acc, &a, 1);
            } else {
                int64_t sa = load_i(cdt, acc), sb = load_i(cdt, v);
                uint64_t r;
                if (op == OP_ADD) {
                    r = (uint64_t)sa + (uint64_t)sb;
                } else if (op == OP_MULTIPLY) {
                    r = (uint64_t)sa * (uint64_t)sb;
                } else {
                    int gt = is_unsigned_int(cdt) ? (uint64_t)sa > (uint64_t)sb : sa > sb;
                    r = (uint64_t)((op == OP_MAXIMUM) == gt ? sa : sb);
                }
                switch (osz) {
                case 1: { uint8_t t = (uint8_t)r; memcpy(acc, &t, 1); break; }
                case 2: { uint16_t t = (uint16_t)r; memcpy(acc, &t, 2); break; }
                case 4: { uint32_t t = (uint32_t)r; memcpy(acc, &t, 4); break; }
                default: memcpy(acc, &r, 8); break;
                }
            }
            memcpy(po, acc, osz);
        }
        for (k = n - 1; k >= 0; k--) {
            if (k == axis) continue;
            if (++idx[k] < in->shape[k]) break;
            idx[k] = 0;
        }
    }
    return ORA_OK;
}

/* argext (ufunc.rs:8131) -> reduce_axis_argext (ufunc.rs:8037): for each
 * kept index (C order), scan the axis in index order; the first NaN wins,
 * otherwise the first strict extreme.  axis < 0 scans the C-order ravel. */
int ora_argext(int ismax, const ora_desc *in, int axis, const ora_desc *out)
{
    int n = in->ndim, k, nk = 0, kept[ORA_MAXDIM], isz = ITEMSIZE[in->dtype], dt = in->dtype;
    int64_t len, nout = 1, o, okidx[ORA_MAXDIM];
    if (axis >= n) return ORA_EAXIS;
    if (axis < 0) {
        len = shape_size(n, in->shape);
        if (len == 0) return ORA_EVALUE;
    } else {
        len = in->shape[axis];
        for (k = 0; k < n; k++) if (k != axis) kept[nk++] = k;
        for (k = 0; k < nk; k++) nout *= in->shape[kept[k]];
        if (len == 0 && nout > 0) return ORA_EVALUE;
    }
    memset(okidx, 0, sizeof okidx);
    for (o = 0; o < nout; o++) {
        int64_t best = 0, i, base = 0, ridx[ORA_MAXDIM];
        double bd = 0;
        int64_t bi = 0;
        uint64_t bu = 0;
        int bnan = 0;
        char *po;
        {
            int64_t ooff = 0;
            for (k = 0; k < nk; k++) ooff += okidx[k] * out->strides[k];
            po = (char *)out->data + ooff * 8;
        }
        for (k = 0; k < nk; k++) base += okidx[k] * in->strides[kept[k]];
        memset(ridx, 0, sizeof ridx);
        for (i = 0; i < len; i++) {
            int64_t off;
            const char *pe;
            if (axis < 0) {
                off = 0;
                for (k = 0; k < n; k++) off += ridx[k] * in->strides[k];
                for (k = n - 1; k >= 0; k--) { if (++ridx[k] < in->shape[k]) break; ridx[k] = 0; }
            } else off = base + i * in->strides[axis];
            pe = (const char *)in->data + off * isz;
            if (is_float_any(dt)) {
                double v = dt == DT_F64 ? *(const double *)pe
                         : dt == DT_F32 ? (double)*(const float *)pe : half_value(*(const uint16_t *)pe);
                int vn = v != v;
                if (i == 0) { bd = v; bnan = vn; best = 0; if (vn) break; continue; }
                if (vn) { best = i; break; }
                if (ismax ? (v > bd) : (v < bd)) { bd = v; best = i; }
            } else if (is_unsigned_int(dt)) {
                uint64_t v = (uint64_t)load_i(dt, pe);
                if (ITEMSIZE[dt] < 8) v &= (1ull << (8 * ITEMSIZE[dt])) - 1;
                if (i == 0 || (ismax ? v > bu : v < bu)) { bu = v; best = i; }
            } else {
                int64_t v = load_i(dt, pe);
                if (i == 0 || (ismax ? v > bi : v < bi)) { bi = v; best = i; }
            }
        }
        (void)bnan;
        *(int64_t *)po = best;
        for (k = nk - 1; k >= 0; k--) {
            if (++okidx[k] < in->shape[kept[k]]) break;
            okidx[k] = 0;
        }
    }
    return ORA_OK;
}

/* Exposed for tests: the pairwise tree of one contiguous run. */
double ora_pairwise_f64(const double *x, int64_t n)
{
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1)), i;
    double r;
    for (i = 0; i < n; i++) off[i] = i * 8;
    r = pw_f64((const char *)x, off, n);
    free(off);
    return r;
}
float ora_pairwise_f32(const float *x, int64_t n)
{
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1)), i;
    float r;
    for (i = 0; i < n; i++) off[i] = i * 4;
    r = pw_f32((const char *)x, off, n);
    free(off);
    return r;
}

/* Exposed for tests: iterator axis order for one operand (innermost first). */
void ora_axis_order(int ndim, const int64_t *shape, const int64_t *strides, int *perm)
{
    int64_t st[1][ORA_MAXDIM];
    int k;
    for (k = 0; k < ndim; k++) st[0][k] = strides[k];
    axis_order(ndim, shape, 1, st, perm);
}
