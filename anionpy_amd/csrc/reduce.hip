// reduce.hip -- gfx950 reduction kernels: add / maximum / minimum .reduce
// along any axes and argmax / argmin.
//
// Replaces the reference's reduce engine: reduce_axis_pairwise_narrow_wide
// (ionp-core/src/ufunc.rs:6658-6799) with pairwise_group / pairwise_sum /
// tree_combine (6371-6531), sequential_group (6546) and reduce_axis_argext
// (8037-8111).
//
// Float add.reduce must equal NumPy's result bit for bit.  NumPy's buffered
// iterator hands its add loop one window of at most `bufsize` (8192)
// elements of the innermost reduced block W at a time, and the loop does
// acc = acc + pairwise(window).  So:
//   pw_windows  - one workgroup per window evaluates NumPy's exact pairwise
//                 tree (8-accumulator 128-element leaves, split points
//                 rounded to multiples of 8) from an LDS copy of the window;
//   pw_fold     - the ordered fold acc = 0; acc += partial[i] per output;
//   pw_narrow   - W empty (reduced axes outside a kept axis): one thread per
//                 output folds its elements in iterator order, with loads
//                 batched ahead of the dependent add chain.
// maximum/minimum: a window's effect on the running value is summarised
// (best value, tie flag, NaN outcome) in parallel and the summaries are
// combined with an associative, order-preserving rule that reproduces the
// x86 AVX-512 reduce loop NumPy runs (lane order, canonical NaN).
// Integer reductions and argmax/argmin are order-free (associative).
#include "reduce.h"

#include <mutex>

#include "elementwise.h"
#include "device_ops.h"

namespace ionp {

// ---------------------------------------------------------------------------
// scratch
// ---------------------------------------------------------------------------
static std::mutex g_scratch_mu;
static void *g_scratch = nullptr;
static size_t g_scratch_bytes = 0;

void *scratch_get(size_t bytes, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (bytes <= g_scratch_bytes) return g_scratch;
    if (g_scratch) {
        (void)hipDeviceSynchronize();  // earlier launches may still read it
        (void)hipFree(g_scratch);
        g_scratch = nullptr;
        g_scratch_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (hipMalloc(&g_scratch, want) != hipSuccess) {
        g_scratch = nullptr;
        return nullptr;
    }
    g_scratch_bytes = want;
    (void)s;
    return g_scratch;
}

// One workgroup per window g = (o, r, w), windows enumerated output-major.
template <typename T>
__global__ __launch_bounds__(kPwThreads) void pw_windows(const RedArgs A, int64_t nwindows, int direct) {
    __shared__ T xs[kPwMaxWin];
    __shared__ T racc[kPwMaxWin / 8];
    __shared__ T val[kPwMaxWin / 8];
    __shared__ int leaf_lo[kPwMaxWin / 64 + 2], leaf_n[kPwMaxWin / 64 + 2];
    __shared__ int cnt[1];
    constexpr int LT = FDT<T>::v;
    for (int64_t g = blockIdx.x; g < nwindows; g += gridDim.x) {
        const int64_t per_out = A.rsize * A.nwin;
        const int64_t o = g / per_out, rem = g - o * per_out;
        const int64_t r = rem / A.nwin, w = rem - r * A.nwin;
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ib += r_offset(A, r);
        const int64_t s0 = w * A.bufsize;
        const int L = (int)min(A.bufsize, A.wsize - s0);
        for (int i = threadIdx.x; i < L; i += blockDim.x) xs[i] = ld<LT>(A.in_dt, A.in, ib + w_offset(A, s0 + i));
        __syncthreads();
        const T res = pairwise_block<T>([&](int i) { return xs[i]; }, L, racc, val, leaf_lo, leaf_n, cnt);
        if (threadIdx.x == 0) {
            if (direct) ((T *)A.out)[ob] = fadd(T(0), res);
            else ((T *)A.scratch)[g] = res;
        }
        __syncthreads();
    }
}

// Ordered fold per output: acc = 0; acc = acc + partial[i] (ufunc.rs:7137
// zero seed; the outer-axis sequential fold of pairwise_group 6504-6531).
template <typename T> __global__ __launch_bounds__(256) void pw_fold(const RedArgs A) {
    const int64_t per_out = A.rsize * A.nwin;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        const T *p = (const T *)A.scratch + o * per_out;
        T acc = T(0);
        int64_t i = 0;
        constexpr int U = 16;
        for (; i + U <= per_out; i += U) {
            T v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = p[i + u];
#pragma unroll
            for (int u = 0; u < U; u++) acc = fadd(acc, v[u]);
        }
        for (; i < per_out; i++) acc = fadd(acc, p[i]);
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((T *)A.out)[ob] = acc;
    }
}

// W empty: per output, acc = 0; acc = acc + x for each reduced element in
// iterator order.  Loads are issued U at a time ahead of the add chain.
template <typename T> __global__ __launch_bounds__(256) void pw_narrow(const RedArgs A) {
    constexpr int LT = FDT<T>::v;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        T acc = T(0);
        constexpr int U = 16;
        int64_t r = 0;
        if (A.nr == 1) {
            const int64_t st = A.r_st[0];
            for (; r + U <= A.rsize; r += U) {
                T v[U];
#pragma unroll
                for (int u = 0; u < U; u++) v[u] = ld<LT>(A.in_dt, A.in, ib + (r + u) * st);
#pragma unroll
                for (int u = 0; u < U; u++) acc = fadd(acc, v[u]);
            }
        }
        for (; r < A.rsize; r++) acc = fadd(acc, ld<LT>(A.in_dt, A.in, ib + r_offset(A, r)));
        ((T *)A.out)[ob] = acc;
    }
}

// ---------------------------------------------------------------------------
// maximum / minimum on floats
// ---------------------------------------------------------------------------
// Effect of a run of windows on the running value p (an associative,
// order-preserving summary):
//   p NaN      -> canonical NaN if any window ran NumPy's SIMD reduce
//                 (`simd`), else p unchanged;
//   otherwise  -> `nanv` if the run produces a NaN, else the best value M
//                 if it beats p, or ties p with `flag` set (its element sits
//                 after every copy of p in the loop's evaluation order).
template <typename T> struct ExtSum {
    T m, nanv;
    uint32_t has_m, flag, has_nan, simd;
};

template <typename T, bool MAX> __device__ inline bool beats(T a, T b) { return MAX ? a > b : a < b; }

template <typename T, bool MAX> __device__ inline ExtSum<T> ext_combine(const ExtSum<T> &a, const ExtSum<T> &b) {
    ExtSum<T> r;
    r.has_nan = a.has_nan | b.has_nan;
    r.nanv = a.has_nan ? (b.simd ? canonical_nan(T(0)) : a.nanv) : b.nanv;
    r.simd = a.simd | b.simd;
    if (!b.has_m || (a.has_m && !(beats<T, MAX>(b.m, a.m) || (b.m == a.m && b.flag)))) {
        r.m = a.m;
        r.flag = a.flag;
        r.has_m = a.has_m;
    } else {
        r.m = b.m;
        r.flag = b.flag;
        r.has_m = b.has_m;
    }
    return r;
}
template <typename T, bool MAX> __device__ inline T ext_apply(const ExtSum<T> &s, T p) {
    if (p != p) return s.simd ? canonical_nan(T(0)) : p;
    if (s.has_nan) return s.nanv;
    if (s.has_m && (beats<T, MAX>(s.m, p) || (s.m == p && s.flag))) return s.m;
    return p;
}

// In-order rank of each SIMD lane in the AVX-512 horizontal reduce of
// NumPy 2.2.6's maximum/minimum loop (later rank wins ties).
__constant__ int kLaneRank64[8] = {3, 7, 1, 5, 2, 6, 0, 4};
__constant__ int kLaneRank32[16] = {3, 7, 11, 15, 1, 5, 9, 13, 2, 6, 10, 14, 0, 4, 8, 12};

// Per-thread candidate within a window: best non-NaN (value, key) with
// later key winning ties, first NaN by key, NaN seen by the vector part.
template <typename T> struct ExtCand {
    T m, nanv;
    int64_t km, kn;
    uint32_t has_m, has_nan, vec_nan;
};

template <typename T, bool MAX> __device__ inline void cand_merge(ExtCand<T> &a, const ExtCand<T> &b) {
    if (b.has_m && (!a.has_m || beats<T, MAX>(b.m, a.m) || (b.m == a.m && b.km > a.km))) {
        a.m = b.m;
        a.km = b.km;
        a.has_m = 1;
    }
    if (b.has_nan && (!a.has_nan || b.kn < a.kn)) {
        a.nanv = b.nanv;
        a.kn = b.kn;
        a.has_nan = 1;
    }
    a.vec_nan |= b.vec_nan;
}

template <typename T, bool MAX>
__global__ __launch_bounds__(256) void ext_windows(const RedArgs A, int64_t nwindows) {
    constexpr int LT = FDT<T>::v;
    constexpr int V = sizeof(T) == 8 ? 8 : 16;
    __shared__ ExtCand<T> red[256];
    for (int64_t g = blockIdx.x; g < nwindows; g += gridDim.x) {
        const int64_t per_out = A.rsize * A.nwin;
        const int64_t o = g / per_out, rem = g - o * per_out;
        const int64_t r = rem / A.nwin, w = rem - r * A.nwin;
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ib += r_offset(A, r);
        const int64_t s0 = w * A.bufsize;
        const int64_t L = min(A.bufsize, A.wsize - s0);
        // a window inside W's first coalesced run is read in place with that
        // run's stride; one spanning runs goes through NumPy's contiguous buffer
        const bool one_run = (s0 / A.w_shape[0]) == ((s0 + L - 1) / A.w_shape[0]);
        const bool simd = one_run ? (A.w_st[0] == 1) : true;
        const int64_t skip = (r == 0 && w == 0) ? 1 : 0;  // first element seeds the fold
        const int64_t len = L - skip;
        const int64_t nv = simd ? len - len % V : (len >= 8 ? len - len % 8 : 0);
        const int64_t per_lane = simd ? nv / V : nv / 8;
        ExtCand<T> c;
        c.has_m = c.has_nan = c.vec_nan = 0;
        c.km = c.kn = 0;
        c.m = c.nanv = T(0);
        for (int64_t e = threadIdx.x; e < len; e += blockDim.x) {
            const T v = ld<LT>(A.in_dt, A.in, ib + w_offset(A, s0 + skip + e));
            int64_t key;
            if (e < nv) {
                const int lane = simd ? (int)(e % V) : (int)(e & 7);
                const int rank = simd ? (V == 8 ? kLaneRank64[lane] : kLaneRank32[lane]) : lane;
                key = (int64_t)rank * per_lane + e / (simd ? V : 8);
            } else {
                key = e;
            }
            if (v != v) {
                if (!c.has_nan || key < c.kn) {
                    c.has_nan = 1;
                    c.kn = key;
                    c.nanv = v;
                }
                if (simd && e < nv) c.vec_nan = 1;
            } else if (!c.has_m || beats<T, MAX>(v, c.m) || (v == c.m && key > c.km)) {
                c.has_m = 1;
                c.m = v;
                c.km = key;
            }
        }
        red[threadIdx.x] = c;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                ExtCand<T> a = red[threadIdx.x];
                cand_merge<T, MAX>(a, red[threadIdx.x + s]);
                red[threadIdx.x] = a;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const ExtCand<T> b = red[0];
            ExtSum<T> S;
            S.simd = simd && len >= 1;
            S.has_m = b.has_m;
            S.m = b.m;
            S.flag = simd ? (b.km >= (int64_t)(V - 1) * per_lane) : 1;
            S.has_nan = b.has_nan;
            S.nanv = (simd && b.vec_nan) ? canonical_nan(T(0)) : b.nanv;
            ((ExtSum<T> *)A.scratch)[g] = S;
        }
        __syncthreads();
    }
}

// Per output: ordered combine of its window summaries (block-wide tree that
// keeps left-to-right order), applied to the seed element.
template <typename T, bool MAX> __global__ __launch_bounds__(256) void ext_fold(const RedArgs A) {
    constexpr int LT = FDT<T>::v;
    __shared__ ExtSum<T> red[256];
    const int64_t per_out = A.rsize * A.nwin;
    for (int64_t o = blockIdx.x; o < A.nout; o += gridDim.x) {
        const ExtSum<T> *S = (const ExtSum<T> *)A.scratch + o * per_out;
        const int64_t per_t = (per_out + blockDim.x - 1) / blockDim.x;
        const int64_t b = threadIdx.x * per_t, e = min(per_out, b + per_t);
        ExtSum<T> acc;
        acc.has_m = acc.has_nan = acc.simd = acc.flag = 0;
        acc.m = acc.nanv = T(0);
        for (int64_t i = b; i < e; i++) acc = ext_combine<T, MAX>(acc, S[i]);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = 1; s < (int)blockDim.x; s <<= 1) {
            if ((threadIdx.x % (2 * s)) == 0 && threadIdx.x + s < blockDim.x)
                red[threadIdx.x] = ext_combine<T, MAX>(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            int64_t ib, ob;
            out_offsets(A, (uint32_t)o, &ib, &ob);
            const T seed = ld<LT>(A.in_dt, A.in, ib);
            ((T *)A.out)[ob] = ext_apply<T, MAX>(red[0], seed);
        }
        __syncthreads();
    }
}

// W empty: acc = first element; acc = maxn(acc, x) in iterator order.
template <typename T, bool MAX> __global__ __launch_bounds__(256) void ext_narrow(const RedArgs A) {
    constexpr int LT = FDT<T>::v;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        T acc = ld<LT>(A.in_dt, A.in, ib);
        for (int64_t r = 1; r < A.rsize; r++) {
            const T v = ld<LT>(A.in_dt, A.in, ib + r_offset(A, r));
            acc = MAX ? fmaxn(acc, v) : fminn(acc, v);
        }
        ((T *)A.out)[ob] = acc;
    }
}

// ---------------------------------------------------------------------------
// order-free reductions: integer / bool add, maximum, minimum
// ---------------------------------------------------------------------------
template <typename AT, int OP> __device__ inline AT assoc(AT a, AT b) {
    if constexpr (OP == IONP_ADD) return (AT)((uint64_t)a + (uint64_t)b);  // two's-complement wrap
    else if constexpr (OP == IONP_MAXIMUM) return a > b ? a : b;
    else return a < b ? a : b;
}

// group element q -> input offset: q = w + wsize * r
__device__ inline int64_t group_offset(const RedArgs &A, int64_t q) {
    if (A.nr == 0) return w_offset(A, q);
    const int64_t r = q / A.wsize, w = q - r * A.wsize;
    return w_offset(A, w) + r_offset(A, r);
}

template <typename AT, int LT> __device__ inline AT load_acc(int dt, const char *base, int64_t off) {
    return (AT)ld<LT>(dt, base, off);
}

// AT: int64_t / uint64_t accumulator; LT: dtype the elements are cast to.
template <typename AT, int LT, int OP>
__global__ __launch_bounds__(256) void int_reduce(const RedArgs A, int64_t seg, int64_t nseg) {
    __shared__ AT red[256];
    using ST = typename dt_type<LT>::T;
    for (int64_t g = blockIdx.x; g < A.nout * nseg; g += gridDim.x) {
        const int64_t o = g / nseg, sg = g - o * nseg;
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t q0 = sg * seg, q1 = min(A.group, q0 + seg);
        AT acc = OP == IONP_ADD ? AT(0) : (AT)load_acc<AT, LT>(A.in_dt, A.in, ib + group_offset(A, q0));
        for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x)
            acc = assoc<AT, OP>(acc, (AT)load_acc<ST, LT>(A.in_dt, A.in, ib + group_offset(A, q)));
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) red[threadIdx.x] = assoc<AT, OP>(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (nseg == 1) ((ST *)A.out)[ob] = (ST)red[0];
            else ((AT *)A.scratch)[g] = red[0];
        }
        __syncthreads();
    }
}
template <typename AT, int LT, int OP> __global__ __launch_bounds__(256) void int_reduce_fold(const RedArgs A, int64_t nseg) {
    using ST = typename dt_type<LT>::T;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        const AT *p = (const AT *)A.scratch + o * nseg;
        AT acc = p[0];
        for (int64_t i = 1; i < nseg; i++) acc = assoc<AT, OP>(acc, p[i]);
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((ST *)A.out)[ob] = (ST)acc;
    }
}
// narrow layouts: one thread per output, elements in any order
template <typename AT, int LT, int OP> __global__ __launch_bounds__(256) void int_reduce_cols(const RedArgs A) {
    using ST = typename dt_type<LT>::T;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        AT acc = OP == IONP_ADD ? AT(0) : (AT)load_acc<ST, LT>(A.in_dt, A.in, ib);
        for (int64_t q = 0; q < A.group; q++)
            acc = assoc<AT, OP>(acc, (AT)load_acc<ST, LT>(A.in_dt, A.in, ib + group_offset(A, q)));
        ((ST *)A.out)[ob] = (ST)acc;
    }
}

// zero-size groups: the identity (add) in every output element
template <typename ST> __global__ void fill_zero(const RedArgs A) {
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((ST *)A.out)[ob] = ST(0);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int grid_cap(int64_t items, int bpc) {
    int64_t g = (int64_t)bpc * num_cus();
    if (g > items) g = items;
    return (int)(g > 0 ? g : 1);
}

static ionp_status launched() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

template <typename AT, int LT, int OP>
static ionp_status run_int(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    // column form when the reduced elements are strided and outputs plenty
    const bool cols = A.nw == 0 || (A.nout >= 16384 && A.group <= 4096);
    if (cols) {
        hipLaunchKernelGGL((int_reduce_cols<AT, LT, OP>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
        return launched();
    }
    int64_t seg = 65536;
    int64_t nseg = (A.group + seg - 1) / seg;
    if (nseg > 1) {
        A.scratch = scratch_get((size_t)(A.nout * nseg) * sizeof(AT), s);
        if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for reduction scratch");
    }
    hipLaunchKernelGGL((int_reduce<AT, LT, OP>), dim3(grid_cap(A.nout * nseg, bpc)), dim3(256), 0, s, A, seg, nseg);
    if (nseg > 1)
        hipLaunchKernelGGL((int_reduce_fold<AT, LT, OP>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A, nseg);
    return launched();
}

template <int LT, int OP> static ionp_status run_int_t(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    if constexpr (LT == IONP_UINT64 || LT == IONP_UINT32 || LT == IONP_UINT16 || LT == IONP_UINT8 || LT == IONP_BOOL)
        return run_int<uint64_t, LT, OP>(A, cfg, s);
    else
        return run_int<int64_t, LT, OP>(A, cfg, s);
}

template <int OP> static ionp_status run_int_dispatch(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    switch (A.out_dt) {
        case IONP_BOOL: return run_int_t<IONP_BOOL, OP>(A, cfg, s);
        case IONP_INT8: return run_int_t<IONP_INT8, OP>(A, cfg, s);
        case IONP_UINT8: return run_int_t<IONP_UINT8, OP>(A, cfg, s);
        case IONP_INT16: return run_int_t<IONP_INT16, OP>(A, cfg, s);
        case IONP_UINT16: return run_int_t<IONP_UINT16, OP>(A, cfg, s);
        case IONP_INT32: return run_int_t<IONP_INT32, OP>(A, cfg, s);
        case IONP_UINT32: return run_int_t<IONP_UINT32, OP>(A, cfg, s);
        case IONP_INT64: return run_int_t<IONP_INT64, OP>(A, cfg, s);
        case IONP_UINT64: return run_int_t<IONP_UINT64, OP>(A, cfg, s);
        default: return set_error(IONP_ERR_NO_LOOP, "no integer reduce loop");
    }
}

template <typename T> static ionp_status run_float_add(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (A.nw == 0) {
        hipLaunchKernelGGL(pw_narrow<T>, dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
        return launched();
    }
    const int64_t nwindows = A.nout * A.rsize * A.nwin;
    const int direct = A.rsize == 1 && A.nwin == 1;
    if (!direct) {
        A.scratch = scratch_get((size_t)nwindows * sizeof(T), s);
        if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for reduction scratch");
    }
    hipLaunchKernelGGL(pw_windows<T>, dim3(grid_cap(nwindows, bpc)), dim3(kPwThreads), 0, s, A, nwindows, direct);
    if (!direct) hipLaunchKernelGGL(pw_fold<T>, dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
    return launched();
}

template <typename T, bool MAX> static ionp_status run_float_ext(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (A.nw == 0) {
        hipLaunchKernelGGL((ext_narrow<T, MAX>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
        return launched();
    }
    const int64_t nwindows = A.nout * A.rsize * A.nwin;
    A.scratch = scratch_get((size_t)nwindows * sizeof(ExtSum<T>), s);
    if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for reduction scratch");
    hipLaunchKernelGGL((ext_windows<T, MAX>), dim3(grid_cap(nwindows, bpc)), dim3(256), 0, s, A, nwindows);
    hipLaunchKernelGGL((ext_fold<T, MAX>), dim3(grid_cap(A.nout, bpc)), dim3(256), 0, s, A);
    return launched();
}

ionp_status run_reduce(const ReducePlan &p, int64_t bufsize, const ionp_config *cfg, hipStream_t s) {
    if (p.nout == 0) return IONP_OK;
    RedArgs A;
    std::memset(&A, 0, sizeof A);
    A.in_dt = p.in_dt;
    A.out_dt = p.out_dt;
    A.op = p.op;
    A.nk = p.nk;
    A.nw = p.nw;
    A.nr = p.nr;
    A.nout = p.nout;
    A.wsize = p.wsize;
    A.rsize = p.rsize;
    A.group = p.group;
    A.bufsize = bufsize;
    A.nwin = p.wsize > 0 ? (p.wsize + bufsize - 1) / bufsize : 0;
    A.in = p.in;
    A.out = p.out;
    if (p.nout >= (1ll << 31) || p.wsize >= (1ll << 31) * 8 || p.rsize >= (1ll << 31))
        return set_error(IONP_ERR_UNSUPPORTED, "reduction extent too large");
    for (int d = 0; d < p.nk; d++) {
        A.k_shape[d] = p.k_shape[d];
        A.k_in_st[d] = p.k_in_st[d];
        A.k_out_st[d] = p.k_out_st[d];
        A.k_div[d] = FastDiv((uint32_t)p.k_shape[d]);
    }
    for (int d = 0; d < p.nw; d++) {
        A.w_shape[d] = p.w_shape[d];
        A.w_st[d] = p.w_st[d];
        A.w_div[d] = FastDiv((uint32_t)(p.w_shape[d] < (1ll << 31) ? p.w_shape[d] : 1));
    }
    if (p.nw > 1 && p.wsize >= (1ll << 31)) return set_error(IONP_ERR_UNSUPPORTED, "reduction extent too large");
    for (int d = 0; d < p.nr; d++) {
        A.r_shape[d] = p.r_shape[d];
        A.r_st[d] = p.r_st[d];
        A.r_div[d] = FastDiv((uint32_t)p.r_shape[d]);
    }
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (p.group == 0) {
        if (p.op != IONP_ADD) return set_error(IONP_ERR_VALUE, std::string("zero-size array to reduction operation ") +
                                                               op_name(p.op) + " which has no identity");
        const int g = grid_cap((p.nout + 255) / 256, bpc);
        switch (itemsize(p.out_dt)) {
            case 1: hipLaunchKernelGGL(fill_zero<uint8_t>, dim3(g), dim3(256), 0, s, A); break;
            case 2: hipLaunchKernelGGL(fill_zero<uint16_t>, dim3(g), dim3(256), 0, s, A); break;
            case 4: hipLaunchKernelGGL(fill_zero<uint32_t>, dim3(g), dim3(256), 0, s, A); break;
            default: hipLaunchKernelGGL(fill_zero<uint64_t>, dim3(g), dim3(256), 0, s, A); break;
        }
        return launched();
    }
    if (is_float_dt(p.out_dt)) {
        if (p.op == IONP_ADD) {
            if (bufsize > kPwMaxWin) return set_error(IONP_ERR_UNSUPPORTED, "bufsize above 8192");
            return p.out_dt == IONP_FLOAT64 ? run_float_add<double>(A, cfg, s) : run_float_add<float>(A, cfg, s);
        }
        if (p.op == IONP_MAXIMUM)
            return p.out_dt == IONP_FLOAT64 ? run_float_ext<double, true>(A, cfg, s) : run_float_ext<float, true>(A, cfg, s);
        return p.out_dt == IONP_FLOAT64 ? run_float_ext<double, false>(A, cfg, s) : run_float_ext<float, false>(A, cfg, s);
    }
    switch (p.op) {
        case IONP_ADD: return run_int_dispatch<IONP_ADD>(A, cfg, s);
        case IONP_MAXIMUM: return run_int_dispatch<IONP_MAXIMUM>(A, cfg, s);
        default: return run_int_dispatch<IONP_MINIMUM>(A, cfg, s);
    }
}

// ---------------------------------------------------------------------------
// argmax / argmin (reduce_axis_argext, ufunc.rs:8037; first NaN wins, then
// the first strict extreme)
// ---------------------------------------------------------------------------
template <typename T, bool MAX> struct ArgBest {
    T v;
    int64_t i;
    uint32_t has, nan;
};
template <typename T, bool MAX> __device__ inline void arg_merge(ArgBest<T, MAX> &a, const ArgBest<T, MAX> &b) {
    if (!b.has) return;
    if (!a.has) { a = b; return; }
    if (a.nan || b.nan) {
        if (b.nan && (!a.nan || b.i < a.i)) a = b;
        return;
    }
    if (beats<T, MAX>(b.v, a.v) || (b.v == a.v && b.i < a.i)) a = b;
}
template <typename T> __device__ inline bool is_nan_any(T v) {
    if constexpr (std::is_floating_point<T>::value) return v != v;
    else return false;
}

__device__ inline void arg_offsets(const ArgArgs &A, uint32_t o, int64_t *ib, int64_t *ob) {
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
    *ib = a;
    *ob = b;
}
__device__ inline int64_t arg_l_offset(const ArgArgs &A, int64_t i) {
    if (A.nl == 1) return i * A.l_st[0];
    int64_t off = 0;
    uint32_t idx = (uint32_t)i;
    for (int d = 0; d < A.nl; d++) {
        uint32_t c;
        if (d + 1 < A.nl) {
            const uint32_t q = A.l_div[d].div(idx);
            c = idx - q * (uint32_t)A.l_shape[d];
            idx = q;
        } else {
            c = idx;
        }
        off += (int64_t)c * A.l_st[d];
    }
    return off;
}

template <typename T, bool MAX> __global__ __launch_bounds__(256) void arg_rows(const ArgArgs A) {
    __shared__ ArgBest<T, MAX> red[256];
    const T *x = (const T *)A.in;
    for (int64_t g = blockIdx.x; g < A.nout * A.nseg; g += gridDim.x) {
        const int64_t o = g / A.nseg, sg = g - o * A.nseg;
        int64_t ib, ob;
        arg_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t i0 = sg * A.seg, i1 = min(A.len, i0 + A.seg);
        ArgBest<T, MAX> b;
        b.has = 0;
        b.nan = 0;
        b.i = 0;
        b.v = T(0);
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            const T v = x[ib + arg_l_offset(A, i)];
            ArgBest<T, MAX> c;
            c.v = v;
            c.i = i;
            c.has = 1;
            c.nan = is_nan_any(v);
            arg_merge(b, c);
        }
        red[threadIdx.x] = b;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                ArgBest<T, MAX> a = red[threadIdx.x];
                arg_merge(a, red[threadIdx.x + s]);
                red[threadIdx.x] = a;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (A.nseg == 1) ((int64_t *)A.out)[ob] = red[0].i;
            else ((ArgBest<T, MAX> *)A.scratch)[g] = red[0];
        }
        __syncthreads();
    }
}
template <typename T, bool MAX> __global__ __launch_bounds__(256) void arg_fold(const ArgArgs A) {
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        const ArgBest<T, MAX> *p = (const ArgBest<T, MAX> *)A.scratch + o * A.nseg;
        ArgBest<T, MAX> b = p[0];
        for (int64_t k = 1; k < A.nseg; k++) arg_merge(b, p[k]);
        int64_t ib, ob;
        arg_offsets(A, (uint32_t)o, &ib, &ob);
        ((int64_t *)A.out)[ob] = b.i;
    }
}
template <typename T, bool MAX> __global__ __launch_bounds__(256) void arg_cols(const ArgArgs A) {
    const T *x = (const T *)A.in;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        arg_offsets(A, (uint32_t)o, &ib, &ob);
        T best = x[ib];
        int64_t bi = 0;
        if (!is_nan_any(best)) {
            for (int64_t i = 1; i < A.len; i++) {
                const T v = x[ib + i * A.l_st[0]];
                if (is_nan_any(v)) { bi = i; break; }
                if (beats<T, MAX>(v, best)) { best = v; bi = i; }
            }
        }
        ((int64_t *)A.out)[ob] = bi;
    }
}

template <typename T, bool MAX> static ionp_status run_arg_t(ArgArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    const bool cols = A.nl == 1 && A.l_st[0] != 1 && A.nout >= 4096 && A.len <= 65536;
    if (cols) {
        hipLaunchKernelGGL((arg_cols<T, MAX>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
        return launched();
    }
    A.seg = 65536;
    A.nseg = (A.len + A.seg - 1) / A.seg;
    if (A.nseg > 1) {
        A.scratch = scratch_get((size_t)(A.nout * A.nseg) * sizeof(ArgBest<T, MAX>), s);
        if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for argext scratch");
    }
    hipLaunchKernelGGL((arg_rows<T, MAX>), dim3(grid_cap(A.nout * A.nseg, bpc)), dim3(256), 0, s, A);
    if (A.nseg > 1) hipLaunchKernelGGL((arg_fold<T, MAX>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
    return launched();
}

template <bool MAX> static ionp_status run_arg_m(ArgArgs &A, const ionp_config *cfg, hipStream_t s) {
    switch (A.in_dt) {
        case IONP_BOOL: return run_arg_t<uint8_t, MAX>(A, cfg, s);
        case IONP_INT8: return run_arg_t<int8_t, MAX>(A, cfg, s);
        case IONP_UINT8: return run_arg_t<uint8_t, MAX>(A, cfg, s);
        case IONP_INT16: return run_arg_t<int16_t, MAX>(A, cfg, s);
        case IONP_UINT16: return run_arg_t<uint16_t, MAX>(A, cfg, s);
        case IONP_INT32: return run_arg_t<int32_t, MAX>(A, cfg, s);
        case IONP_UINT32: return run_arg_t<uint32_t, MAX>(A, cfg, s);
        case IONP_INT64: return run_arg_t<int64_t, MAX>(A, cfg, s);
        case IONP_UINT64: return run_arg_t<uint64_t, MAX>(A, cfg, s);
        case IONP_FLOAT32: return run_arg_t<float, MAX>(A, cfg, s);
        default: return run_arg_t<double, MAX>(A, cfg, s);
    }
}

ionp_status run_argext(ArgArgs &A, const ionp_config *cfg, hipStream_t s) {
    if (A.nout == 0) return IONP_OK;
    if (A.nout >= (1ll << 31) || (A.nl > 1 && A.len >= (1ll << 31)))
        return set_error(IONP_ERR_UNSUPPORTED, "argext extent too large");
    for (int d = 0; d < A.nk; d++) A.k_div[d] = FastDiv((uint32_t)A.k_shape[d]);
    for (int d = 0; d < A.nl; d++) A.l_div[d] = FastDiv((uint32_t)(A.l_shape[d] < (1ll << 31) ? A.l_shape[d] : 1));
    return A.ismax ? run_arg_m<true>(A, cfg, s) : run_arg_m<false>(A, cfg, s);
}

}  // namespace ionp
