// accumulate.hip -- ufunc.accumulate for add / multiply / maximum / minimum
// (accumulate_axis, ionp-core/src/ufunc.rs:7967 -> accumulate_axis_generic
// 7830): along one axis, out[0] = x[0] and out[i] = op(out[i-1], x[i]) in
// the output dtype.  The scan is sequential per fiber (NumPy's float add
// accumulate is the plain running sum, no pairwise tree), so the GPU runs
// fibers in parallel:
//   acc_thread - a thread per fiber, loads issued U ahead of the chain
//                (fibers adjacent in memory);
//   acc_tiled  - fibers contiguous along the axis: 16 per wavefront,
//                chunks of 64 elements transposed through LDS;
//   acc_wave   - few long fibers: a wavefront per fiber stages blocks in
//                LDS and one lane runs the chain over them;
//   acc_block  - many fibers: a workgroup stages F fibers x R steps in LDS,
//                F lanes run the chains, the next block's loads in flight.
// Float add and multiply run plain IEEE arithmetic and redo a fiber that
// ends in NaN with fadd's / fmul's x86 NaN rules (they agree until a NaN
// appears, and a NaN stays).
// An input of another dtype is first cast into scratch memory, as the
// reference casts before scanning.
#include "accumulate.h"

#include "device_ops.h"
#include "elementwise.h"

namespace ionp {

struct AccArgs {
    int64_t len;              // extent of the scanned axis
    int64_t in_st, out_st;    // strides along it (elements)
    int64_t nfib;             // fibers (product of the other extents)
    int32_t nf;               // fiber dims, innermost first
    int64_t fshape[IONP_MAX_DIMS];
    FastDiv fdiv[IONP_MAX_DIMS];
    int64_t fin[IONP_MAX_DIMS], fout[IONP_MAX_DIMS];
    const char *in;
    char *out;
};

__device__ inline void fiber_offsets(const AccArgs &A, uint32_t f, int64_t *ib, int64_t *ob) {
    int64_t a = 0, b = 0;
    for (int d = 0; d < A.nf; d++) {
        uint32_t c;
        if (d + 1 < A.nf) {
            const uint32_t q = A.fdiv[d].div(f);
            c = f - q * (uint32_t)A.fshape[d];
            f = q;
        } else {
            c = f;
        }
        a += (int64_t)c * A.fin[d];
        b += (int64_t)c * A.fout[d];
    }
    *ib = a;
    *ob = b;
}

// one step of the scan in dtype LT (float add / multiply: plain IEEE, see
// above; bool: add and maximum are or, multiply and minimum and)
template <int LT, int OP> __device__ inline typename dt_type<LT>::T acc_step(typename dt_type<LT>::T a,
                                                                            typename dt_type<LT>::T b) {
    using T = typename dt_type<LT>::T;
    if constexpr (LT == IONP_BOOL) {
        return (OP == IONP_MINIMUM || OP == IONP_MULTIPLY) ? (T)((a != 0) & (b != 0)) : (T)((a != 0) | (b != 0));
    } else if constexpr (std::is_floating_point<T>::value) {
        if constexpr (OP == IONP_ADD) return a + b;
        else if constexpr (OP == IONP_MULTIPLY) return a * b;
        else if constexpr (OP == IONP_MAXIMUM) return fmaxn(a, b);
        else return fminn(a, b);
    } else {
        if constexpr (OP == IONP_ADD) return (T)((uint64_t)a + (uint64_t)b);  // two's-complement wrap
        else if constexpr (OP == IONP_MULTIPLY) return (T)((uint64_t)a * (uint64_t)b);
        else if constexpr (OP == IONP_MAXIMUM) return a > b ? a : b;
        else return a < b ? a : b;
    }
}

// float add and multiply: the plain IEEE chain needs the x86 NaN redo
template <typename T, int OP> constexpr bool acc_plain() {
    return std::is_floating_point<T>::value && (OP == IONP_ADD || OP == IONP_MULTIPLY);
}
template <int LT, int OP> __device__ inline bool acc_redo(typename dt_type<LT>::T acc) {
    if constexpr (acc_plain<typename dt_type<LT>::T, OP>()) return acc != acc;
    else return false;
}

// the fiber again with fadd / fmul (x86 NaN rules), every output rewritten
template <typename T, int OP>
__device__ inline void acc_redo_fiber(const AccArgs &A, const T *x, T *o) {
    T acc = x[0];
    o[0] = acc;
    for (int64_t i = 1; i < A.len; i++) {
        acc = OP == IONP_MULTIPLY ? fmul(acc, x[i * A.in_st]) : fadd(acc, x[i * A.in_st]);
        o[i * A.out_st] = acc;
    }
}

template <int LT, int OP> __global__ __launch_bounds__(256) void acc_thread(const AccArgs A) {
    using T = typename dt_type<LT>::T;
    constexpr int U = 16;
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < A.nfib; f += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        fiber_offsets(A, (uint32_t)f, &ib, &ob);
        const T *x = (const T *)A.in + ib;
        T *o = (T *)A.out + ob;
        T acc = x[0];
        o[0] = acc;
        int64_t i = 1;
        // double-buffered: the next U loads are in flight while U steps run
        T cur[U], nxt[U];
        if (i + U <= A.len) {
#pragma unroll
            for (int u = 0; u < U; u++) cur[u] = x[(i + u) * A.in_st];
        }
        for (; i + U <= A.len; i += U) {
            if (i + 2 * U <= A.len) {
#pragma unroll
                for (int u = 0; u < U; u++) nxt[u] = x[(i + U + u) * A.in_st];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                acc = acc_step<LT, OP>(acc, cur[u]);
                o[(i + u) * A.out_st] = acc;
            }
#pragma unroll
            for (int u = 0; u < U; u++) cur[u] = nxt[u];
        }
        for (; i < A.len; i++) {
            acc = acc_step<LT, OP>(acc, x[i * A.in_st]);
            o[i * A.out_st] = acc;
        }
        if constexpr (acc_plain<T, OP>())
            if (acc_redo<LT, OP>(acc)) acc_redo_fiber<T, OP>(A, x, o);
    }
}

// Few long fibers: one wavefront per fiber.  Blocks of 1024 elements are
// loaded by all 64 lanes into LDS (16 each, all in flight), lane 0 runs the
// chain over the block in LDS (reads issued 16 ahead of the dependent op,
// results written back in place), and the lanes store the block.
template <int LT, int OP> __global__ __launch_bounds__(64) void acc_wave(const AccArgs A) {
    using T = typename dt_type<LT>::T;
    constexpr int Q = 16, BLK = 64 * Q;
    __shared__ T buf[BLK];
    const int lane = threadIdx.x;
    for (int64_t f = blockIdx.x; f < A.nfib; f += gridDim.x) {
        int64_t ib, ob;
        fiber_offsets(A, (uint32_t)f, &ib, &ob);
        const T *x = (const T *)A.in + ib;
        T *o = (T *)A.out + ob;
        const int64_t n = A.len;
        T acc = T(0);
        for (int64_t rb = 0; rb < n; rb += BLK) {
            const int cnt = (int)min((int64_t)BLK, n - rb);
            T v[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int e = q * 64 + lane;
                v[q] = e < cnt ? x[(rb + e) * A.in_st] : T(0);
            }
#pragma unroll
            for (int q = 0; q < Q; q++) buf[q * 64 + lane] = v[q];
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (lane == 0) {
                int e = 0;
                if (rb == 0) {  // out[0] = x[0]
                    acc = buf[0];
                    e = 1;
                }
                for (; e + 16 <= cnt; e += 16) {
                    T t[16];
#pragma unroll
                    for (int u = 0; u < 16; u++) t[u] = buf[e + u];
#pragma unroll
                    for (int u = 0; u < 16; u++) {
                        acc = acc_step<LT, OP>(acc, t[u]);
                        buf[e + u] = acc;
                    }
                }
                for (; e < cnt; e++) {
                    acc = acc_step<LT, OP>(acc, buf[e]);
                    buf[e] = acc;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int e = q * 64 + lane;
                if (e < cnt) o[(rb + e) * A.out_st] = buf[e];
            }
            __builtin_amdgcn_wave_barrier();
        }
        acc = readlane_t(acc, 0);
        if constexpr (acc_plain<T, OP>())
            if (acc_redo<LT, OP>(acc)) {
                __threadfence();  // the wavefront's plain stores land before lane 0 rewrites them
                if (lane == 0) acc_redo_fiber<T, OP>(A, x, o);
            }
    }
}

// Many fibers contiguous along the axis: a wavefront takes FPW fibers and
// walks them in chunks of 64 elements.  A chunk is loaded with the lanes
// along the axis (every load 64 adjacent elements), transposed through LDS,
// each lane scans its own fiber's 64 elements (running value carried across
// chunks), and the chunk goes back through LDS to coalesced stores.
constexpr int kAccFpw = 16;  // fibers per wavefront in acc_tiled

template <int LT, int OP> __global__ __launch_bounds__(64) void acc_tiled(const AccArgs A) {
    using T = typename dt_type<LT>::T;
    constexpr int FPW = kAccFpw;
    __shared__ T tile[FPW][65];  // [fiber][element], padded against bank conflicts
    const int lane = threadIdx.x;
    const int64_t ngroups = (A.nfib + FPW - 1) / FPW;
    for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int64_t f0 = g * FPW;
        const int nf = (int)min((int64_t)FPW, A.nfib - f0);
        // lane r < nf: fiber f0 + r (its scan, and its base offsets for the copies)
        int64_t myib = 0, myob = 0;
        if (lane < nf) fiber_offsets(A, (uint32_t)(f0 + lane), &myib, &myob);
        int64_t ibr[FPW], obr[FPW];
#pragma unroll
        for (int r = 0; r < FPW; r++) {
            ibr[r] = __shfl(myib, r, 64);
            obr[r] = __shfl(myob, r, 64);
        }
        T acc = T(0);
        T v[FPW], nv[FPW];
        auto load = [&](int64_t c0, T (&dst)[FPW]) {
            const int m = (int)min((int64_t)64, A.len - c0);
#pragma unroll
            for (int r = 0; r < FPW; r++)
                dst[r] = (r < nf && lane < m) ? ((const T *)A.in)[ibr[r] + (c0 + lane) * A.in_st] : T(0);
        };
        load(0, v);
        for (int64_t c0 = 0; c0 < A.len; c0 += 64) {
            const int m = (int)min((int64_t)64, A.len - c0);
#pragma unroll
            for (int r = 0; r < FPW; r++) tile[r][lane] = v[r];
            if (c0 + 64 < A.len) load(c0 + 64, nv);  // next chunk in flight during the scan
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (lane < nf) {
                int e = 0;
                if (c0 == 0) {
                    acc = tile[lane][0];
                    e = 1;
                }
                for (; e < m; e++) {
                    acc = acc_step<LT, OP>(acc, tile[lane][e]);
                    tile[lane][e] = acc;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < FPW; r++)
                if (r < nf && lane < m) ((T *)A.out)[obr[r] + (c0 + lane) * A.out_st] = tile[r][lane];
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int r = 0; r < FPW; r++) v[r] = nv[r];
        }
        if constexpr (acc_plain<T, OP>()) {
            const bool redo = lane < nf && acc != acc;
            if (__ballot(redo)) {
                __threadfence();  // the plain stores land before they are rewritten
                if (redo) acc_redo_fiber<T, OP>(A, (const T *)A.in + myib, (T *)A.out + myob);
            }
        }
    }
}

// Many fibers, blocked: a 256-thread workgroup owns F fibers and walks the
// axis in blocks of R steps.  All threads load a block (F x R elements,
// nontemporal) into LDS -- AX1 (fibers contiguous along the axis): each
// load instruction covers consecutive steps of one fiber; otherwise
// (neighbouring fibers adjacent in memory) consecutive fibers of one step
// -- F lanes each run their fiber's sequential chain over the block in LDS
// (NumPy's order, one dependent op per step), and all threads store the
// block.  The next block's loads are in flight during the chain.  Measured
// on MI355X, 8192 x 8192 cumsum: float64 axis 0 (F 16, R 128) 204-210 us,
// axis 1 (F 16, R 256) 210-215 us (5.0-5.3 TB/s); float32 axis 0 163 us,
// axis 1 135 us (with a guard on each load and store: 170 / 148 us, the
// compiler then waiting for the previous block's stores on every block);
// a thread per fiber, or 16 fibers per wavefront in 64-element chunks, kept
// too few bytes in flight (float64 685 / 529 us).  Neither two blocks of
// loads in flight nor two LDS buffers (stores drained during the next
// chain) ran faster.
template <int LT, int OP, int F, int R, bool AX1> __global__ __launch_bounds__(256) void acc_block(const AccArgs A) {
    using T = typename dt_type<LT>::T;
    constexpr int PER = F * R / 256;  // elements per thread per block
    static_assert(F * R % 256 == 0 && F <= 256, "block shape");
    __shared__ T buf[AX1 ? F : R][(AX1 ? R : F) + 1];  // [fiber][step] or [step][fiber], padded
    __shared__ int64_t sib[F], sob[F];
    const int tid = threadIdx.x;
    // element e of this thread in a block: fiber fl(e), step st(e)
    auto fl = [&](int e) { const int t = tid + 256 * e; return AX1 ? t / R : t % F; };
    auto st = [&](int e) { const int t = tid + 256 * e; return AX1 ? t % R : t / F; };
    auto lds = [&](int f, int k) -> T & {
        if constexpr (AX1) return buf[f][k];
        else return buf[k][f];
    };
    const int64_t ngroups = (A.nfib + F - 1) / F;
    for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const int64_t f0 = g * F;
        const int nf = (int)min((int64_t)F, A.nfib - f0);
        if (tid < F) {
            int64_t ib = 0, ob = 0;
            if (tid < nf) fiber_offsets(A, (uint32_t)(f0 + tid), &ib, &ob);
            sib[tid] = ib;
            sob[tid] = ob;
        }
        __syncthreads();
        const T *x = (const T *)A.in;
        T *o = (T *)A.out;
        // Every load and store is issued unguarded, so the compiler's wait
        // for a block's loads can count exactly and leave the previous
        // block's stores in flight (behind a guard that may skip them it
        // waits for everything, the stores' write latency included, on every
        // block).  An element outside the group's fibers or past the axis
        // end uses the nearest one inside instead: its load reads a valid
        // address and is never used, its store rewrites that element's own
        // value at its own address.
        int fc[PER];  // this thread's fibers, clamped to the group
        int64_t ib[PER], ob[PER];
#pragma unroll
        for (int e = 0; e < PER; e++) {
            fc[e] = min(fl(e), nf - 1);
            ib[e] = sib[fc[e]];
            ob[e] = sob[fc[e]];
        }
        T regs[PER];
        auto load = [&](int64_t i0) {  // block i0 (steps past the end: the last step)
            const int64_t lim = A.len - 1 - i0;
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int64_t k = min((int64_t)st(e), lim);
                regs[e] = __builtin_nontemporal_load(x + ib[e] + (i0 + k) * A.in_st);
            }
        };
        auto stage = [&]() {
#pragma unroll
            for (int e = 0; e < PER; e++) lds(fl(e), st(e)) = regs[e];
        };
        T acc = T(0);
        load(0);
        stage();
        // per block of R steps, staged in LDS: the next block's loads are
        // issued, the chains run, the block is stored, the next is staged
        for (int64_t i0 = 0; i0 < A.len; i0 += R) {
            const int m = (int)min((int64_t)R, A.len - i0);
            load(i0 + R);  // past the end: rereads the last step, unused
            __syncthreads();
            if (tid < nf) {  // the chain of fiber `tid` over this block
                int k = 0;
                if (i0 == 0) {  // out[0] = x[0]
                    acc = lds(tid, 0);
                    k = 1;
                }
                for (; k + 8 <= m; k += 8) {
                    T t[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) t[u] = lds(tid, k + u);
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        acc = acc_step<LT, OP>(acc, t[u]);
                        lds(tid, k + u) = acc;
                    }
                }
                for (; k < m; k++) {
                    acc = acc_step<LT, OP>(acc, lds(tid, k));
                    lds(tid, k) = acc;
                }
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < PER; e++) {
                const int k = min(st(e), m - 1);
                __builtin_nontemporal_store(lds(fc[e], k), o + ob[e] + (i0 + k) * A.out_st);
            }
            __syncthreads();  // also orders every wave's stores before a fiber's redo below
            stage();
        }
        if constexpr (acc_plain<T, OP>()) {
            if (tid < nf && acc != acc) {
                __threadfence();  // the block's plain stores land before they are rewritten
                acc_redo_fiber<T, OP>(A, x + sib[tid], o + sob[tid]);
            }
        }
        __syncthreads();  // sib / sob are rewritten for the next group
    }
}

// float16 scans: each step rounded to float16 (hstep), a thread per fiber
__global__ __launch_bounds__(256) void acc_half(const AccArgs A, int op) {
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < A.nfib; f += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        fiber_offsets(A, (uint32_t)f, &ib, &ob);
        const uint16_t *x = (const uint16_t *)A.in + ib;
        uint16_t *o = (uint16_t *)A.out + ob;
        uint16_t acc = x[0];
        o[0] = acc;
        for (int64_t i = 1; i < A.len; i++) {
            acc = hstep(op, acc, x[i * A.in_st]);
            o[i * A.out_st] = acc;
        }
    }
}

typedef void (*AccFn)(const AccArgs);

enum { ACC_THREAD = 0, ACC_WAVE = 1, ACC_TILED = 2, ACC_BLOCK_AX = 3, ACC_BLOCK_ADJ = 4 };
constexpr int kAccBlockF = 16;  // fibers per acc_block workgroup

// fibers per acc_block workgroup: 8, 16 (default) or 32 (ionp_config.split)
template <int LT, int OP> static AccFn acc_pick(int kind, int fpb) {
    switch (kind) {
        case ACC_WAVE: return acc_wave<LT, OP>;
        case ACC_TILED: return acc_tiled<LT, OP>;
        case ACC_BLOCK_AX:
            return fpb == 8 ? acc_block<LT, OP, 8, 256, true>
                 : fpb == 32 ? acc_block<LT, OP, 32, 128, true> : acc_block<LT, OP, 16, 256, true>;
        case ACC_BLOCK_ADJ:
            return fpb == 8 ? acc_block<LT, OP, 8, 256, false>
                 : fpb == 32 ? acc_block<LT, OP, 32, 64, false> : acc_block<LT, OP, 16, 128, false>;
        default: return acc_thread<LT, OP>;
    }
}

template <int LT> static AccFn acc_fn(int op, int kind, int fpb) {
    switch (op) {
        case IONP_ADD: return acc_pick<LT, IONP_ADD>(kind, fpb);
        case IONP_MULTIPLY: return acc_pick<LT, IONP_MULTIPLY>(kind, fpb);
        case IONP_MAXIMUM: return acc_pick<LT, IONP_MAXIMUM>(kind, fpb);
        default: return acc_pick<LT, IONP_MINIMUM>(kind, fpb);
    }
}

static AccFn acc_kernel(int dt, int op, int wave, int fpb) {
    switch (dt) {
        case IONP_BOOL: return acc_fn<IONP_BOOL>(op, wave, fpb);
        case IONP_INT8: return acc_fn<IONP_INT8>(op, wave, fpb);
        case IONP_UINT8: return acc_fn<IONP_UINT8>(op, wave, fpb);
        case IONP_INT16: return acc_fn<IONP_INT16>(op, wave, fpb);
        case IONP_UINT16: return acc_fn<IONP_UINT16>(op, wave, fpb);
        case IONP_INT32: return acc_fn<IONP_INT32>(op, wave, fpb);
        case IONP_UINT32: return acc_fn<IONP_UINT32>(op, wave, fpb);
        case IONP_INT64: return acc_fn<IONP_INT64>(op, wave, fpb);
        case IONP_UINT64: return acc_fn<IONP_UINT64>(op, wave, fpb);
        case IONP_FLOAT32: return acc_fn<IONP_FLOAT32>(op, wave, fpb);
        default: return acc_fn<IONP_FLOAT64>(op, wave, fpb);
    }
}

// in (already in the output dtype) -> out along `axis`
ionp_status run_accumulate(int op, const ionp_array_desc *in, int axis, const ionp_array_desc *out,
                           const ionp_config *cfg, hipStream_t s) {
    const int n = in->ndim;
    AccArgs A;
    std::memset(&A, 0, sizeof A);
    A.len = in->shape[axis];
    A.in_st = in->strides[axis];
    A.out_st = out->strides[axis];
    // fiber dims ordered innermost first by the operands' strides, so
    // neighbouring threads touch neighbouring memory
    int64_t st[2][IONP_MAX_DIMS];
    std::memcpy(st[0], in->strides, sizeof(int64_t) * n);
    std::memcpy(st[1], out->strides, sizeof(int64_t) * n);
    int perm[IONP_MAX_DIMS];
    iter_axis_order(n, in->shape, 2, st, perm);
    A.nfib = 1;
    for (int k = 0; k < n; k++) {
        const int d = perm[k];
        if (d == axis || in->shape[d] == 1) continue;
        A.fshape[A.nf] = in->shape[d];
        A.fdiv[A.nf] = FastDiv((uint32_t)in->shape[d]);
        A.fin[A.nf] = in->strides[d];
        A.fout[A.nf] = out->strides[d];
        A.nf++;
        A.nfib *= in->shape[d];
    }
    if (A.len == 0 || A.nfib == 0) return IONP_OK;
    if (A.nfib >= (1ll << 31)) return set_error(IONP_ERR_UNSUPPORTED, "accumulate with 2^31 fibers or more");
    const int esz = itemsize(out->dtype);
    A.in = (const char *)in->data + in->offset * esz;
    A.out = (char *)out->data + out->offset * esz;
    // contiguous along the axis: 64 fibers per wavefront through LDS when
    // there are enough of them, else the LDS chain per fiber; otherwise
    // (neighbouring fibers adjacent in memory) a thread per fiber
    if (out->dtype == IONP_FLOAT16) {
        const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
        int64_t grid = (int64_t)bpc * num_cus(), units = (A.nfib + 255) / 256;
        if (grid > units) grid = units;
        hipLaunchKernelGGL(acc_half, dim3((unsigned)grid), dim3(256), 0, s, A, op);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, hipGetErrorString(e));
    }
    // Many fibers with long enough axes: the blocked kernel (fibers
    // contiguous along the axis, or neighbouring fibers adjacent); else 16
    // fibers per wavefront for contiguous fibers, a wavefront per fiber for
    // a few long ones, a thread per fiber otherwise
    int kind;
    const bool adjacent = A.nf > 0 && A.fin[0] == 1 && A.fout[0] == 1;
    if (A.nfib >= 2048 && A.len >= 128 && A.in_st == 1) kind = ACC_BLOCK_AX;
    else if (A.nfib >= 2048 && A.len >= 64 && adjacent) kind = ACC_BLOCK_ADJ;
    else if (A.in_st == 1 && A.nfib >= 256) kind = ACC_TILED;
    else if (A.nfib < 4096 && A.len >= 256) kind = ACC_WAVE;
    else kind = ACC_THREAD;
    const bool block = kind == ACC_BLOCK_AX || kind == ACC_BLOCK_ADJ;
    const int fpb = (cfg && (cfg->split == 8 || cfg->split == 32)) ? cfg->split : kAccBlockF;
    const AccFn fn = acc_kernel(out->dtype, op, kind, fpb);
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    const int threads = kind == ACC_THREAD || block ? 256 : 64;
    const int64_t units = kind == ACC_THREAD ? (A.nfib + 255) / 256
                          : kind == ACC_TILED ? (A.nfib + kAccFpw - 1) / kAccFpw
                          : block             ? (A.nfib + fpb - 1) / fpb
                                              : A.nfib;
    int64_t grid = (int64_t)bpc * num_cus() * (threads == 64 ? 4 : 1);
    if (grid > units) grid = units;
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(threads), 0, s, A);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, hipGetErrorString(e));
}

}  // namespace ionp
