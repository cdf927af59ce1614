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

// One output element of the generic kernel, operand addresses resolved.
template <int LT>
__device__ inline void ew_element(const EwArgs &A, const char *p0, const char *p1, const char *p2, char *po) {
    using T = typename dt_type<LT>::T;
    if (A.op == OP_CAST) {
        store_as<T>(A.out_dt, po, load_as<LT>(A.in_dt[0], p0));
    } else if (A.op == OP_WHERE) {
        const bool c = load_as<IONP_BOOL>(A.in_dt[0], p0) != 0;
        store_as<T>(A.out_dt, po, c ? load_as<LT>(A.in_dt[1], p1) : load_as<LT>(A.in_dt[2], p2));
    } else if (A.op >= IONP_LESS) {
        bool r;
        if (A.mixed_su) {
            r = A.in_dt[0] == IONP_INT64 ? cmp_su(A.op, *(const int64_t *)p0, *(const uint64_t *)p1)
                                         : cmp_su(swap_cmp(A.op), *(const int64_t *)p1, *(const uint64_t *)p0);
        } else {
            r = cmp_op<T>(A.op, load_as<LT>(A.in_dt[0], p0), load_as<LT>(A.in_dt[1], p1));
        }
        *(uint8_t *)po = (uint8_t)r;
    } else {
        store_as<T>(A.out_dt, po, arith_op<LT>(A.op, load_as<LT>(A.in_dt[0], p0), load_as<LT>(A.in_dt[1], p1)));
    }
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
        default: return cvt<LT>(((const double *)base)[off]);
    }
}

template <typename T> struct FDT;
template <> struct FDT<float> { static constexpr int v = IONP_FLOAT32; };
template <> struct FDT<double> { static constexpr int v = IONP_FLOAT64; };

// ---------------------------------------------------------------------------
// NumPy's pairwise tree over an LDS window (ufunc.rs:6416-6448)
// ---------------------------------------------------------------------------
// Node split of pairwise_sum: n2 = n/2 rounded down to a multiple of 8.
__device__ inline int pw_split(int n) { return (n >> 1) & ~7; }

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
                T r = T(0);
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
