// batch.hip -- many small uneven ufunc / reduce calls in one launch.
// No counterpart in the reference (it executes one call at a time).
//
// ionp_batch_create plans every call on the host once (the same planner the
// single-call entry points use), cuts each call into tiles and uploads one
// call table.  ionp_batch_run then executes the whole batch as:
//   PERSISTENT - one launch of a persistent kernel: workgroups pull tile
//                indices from an atomic counter, find the owning call by
//                binary search over the tile prefix, and run the tile with
//                the same device code as the single-call kernels.  Reductions
//                whose outputs span several tiles write partials, which a
//                second small launch folds in NumPy's order.
//   GRAPH      - the single-call launch sequence of every call, captured
//                once into a hipGraph and replayed.
//   PER_CALL   - the single-call entry points, one call after another.
// All three produce bit-identical outputs.
#include <vector>

#include "device_ops.h"

namespace ionp {

enum : int32_t {
    BK_ELEMENTWISE = 0,  // binary / where / cast
    BK_FADD_WIDE = 1,    // float add.reduce, windows over W
    BK_FADD_NARROW = 2,  // float add.reduce, W empty
    BK_IADD = 3,         // integer / bool add.reduce (order-free)
    BK_FILL_ZERO = 4,    // empty reduction groups
};

constexpr int kBatchThreads = 256;
constexpr int64_t kEwTile = kBatchThreads * 16;  // elements per elementwise tile
constexpr int64_t kIntSeg = 8192;                // elements per integer-reduce tile
constexpr int64_t kNarrowOutputs = kBatchThreads;  // outputs per narrow tile

struct BatchCall {
    int32_t kind;
    int32_t direct;    // reduce result written by the tile itself (one tile per output)
    int64_t tiles;     // tiles of this call
    int64_t nseg;      // integer reduce: tiles per output
    int64_t part_off;  // byte offset of this call's partials
    EwArgs ew;
    RedArgs red;
};

}  // namespace ionp

struct ionp_batch {
    int32_t n;
    int64_t ntiles, bytes, part_bytes;
    ionp::BatchCall *d_calls;   // device call table
    int64_t *d_tile_begin;      // device: first tile of each call (prefix), n + 1 entries
    uint32_t *d_counters;       // device: [0] tile counter
    int32_t *d_fold_list;       // device: calls whose partials need a fold
    int32_t nfold;
    char *d_partials;           // device: reduction partials
    std::vector<ionp_call> calls;  // host copy for GRAPH / PER_CALL modes
    hipGraphExec_t graph;
    bool graph_ready;
    size_t scratch_need;
};

namespace ionp {

__global__ void batch_zero(uint32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

template <int LT> __device__ inline void ew_range(const EwArgs &A, uint32_t e0, uint32_t e1) {
    const int i0 = itemsize(A.in_dt[0]), i1 = itemsize(A.in_dt[1]), i2 = itemsize(A.in_dt[2]);
    const int io = itemsize(A.out_dt);
    for (uint32_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        uint32_t rem = e;
        int64_t o0 = 0, o1 = 0, o2 = 0, oo = 0;
        for (int d = 0; d < A.ndim; d++) {
            uint32_t c;
            if (d + 1 < A.ndim) {
                const uint32_t q = A.div[d].div(rem);
                c = rem - q * (uint32_t)A.shape[d];
                rem = q;
            } else {
                c = rem;
            }
            o0 += (int64_t)c * A.in_st[0][d];
            o1 += (int64_t)c * A.in_st[1][d];
            o2 += (int64_t)c * A.in_st[2][d];
            oo += (int64_t)c * A.out_st[d];
        }
        ew_element<LT>(A, A.in[0] + o0 * i0, A.in[1] + o1 * i1, A.in[2] + o2 * i2, A.out + oo * io);
    }
}

__device__ void ew_tile(const EwArgs &A, int64_t t) {
    const int64_t e0 = t * kEwTile, e1 = min(A.total, e0 + kEwTile);
    switch (A.loop_dt) {
        case IONP_BOOL: ew_range<IONP_BOOL>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_INT8: ew_range<IONP_INT8>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_UINT8: ew_range<IONP_UINT8>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_INT16: ew_range<IONP_INT16>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_UINT16: ew_range<IONP_UINT16>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_INT32: ew_range<IONP_INT32>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_UINT32: ew_range<IONP_UINT32>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_INT64: ew_range<IONP_INT64>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_UINT64: ew_range<IONP_UINT64>(A, (uint32_t)e0, (uint32_t)e1); break;
        case IONP_FLOAT32: ew_range<IONP_FLOAT32>(A, (uint32_t)e0, (uint32_t)e1); break;
        default: ew_range<IONP_FLOAT64>(A, (uint32_t)e0, (uint32_t)e1); break;
    }
}

struct PwScratch {
    double racc[kPwMaxWin / 8];
    double val[kPwMaxWin / 8];
    int leaf_lo[kPwMaxWin / 64 + 2], leaf_n[kPwMaxWin / 64 + 2];
    int cnt[1];
};

// one buffer window of a wide float add.reduce (tile g = window index)
template <typename T> __device__ void fadd_window(const BatchCall &C, char *partials, int64_t g, PwScratch &S) {
    constexpr int LT = FDT<T>::v;
    const RedArgs &A = C.red;
    const int64_t per_out = A.rsize * A.nwin;
    const int64_t o = g / per_out, rem = g - o * per_out;
    const int64_t r = rem / A.nwin, w = rem - r * A.nwin;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    ib += r_offset(A, r);
    const int64_t s0 = w * A.bufsize;
    const int L = (int)min(A.bufsize, A.wsize - s0);
    const T res = pairwise_block<T>([&](int i) { return ld<LT>(A.in_dt, A.in, ib + w_offset(A, s0 + i)); }, L,
                                    (T *)S.racc, (T *)S.val, S.leaf_lo, S.leaf_n, S.cnt);
    if (threadIdx.x == 0) {
        if (C.direct) ((T *)A.out)[ob] = fadd(T(0), res);
        else ((T *)(partials + C.part_off))[g] = res;
    }
}

template <typename T> __device__ void fadd_fold(const BatchCall &C, const char *partials) {
    const RedArgs &A = C.red;
    const int64_t per_out = A.rsize * A.nwin;
    for (int64_t o = threadIdx.x; o < A.nout; o += blockDim.x) {
        const T *p = (const T *)(partials + C.part_off) + o * per_out;
        T acc = T(0);
        for (int64_t i = 0; i < per_out; i++) acc = fadd(acc, p[i]);
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((T *)A.out)[ob] = acc;
    }
}

template <typename T> __device__ void fadd_narrow(const BatchCall &C, int64_t t) {
    constexpr int LT = FDT<T>::v;
    const RedArgs &A = C.red;
    const int64_t o = t * kNarrowOutputs + threadIdx.x;
    if (o >= A.nout) return;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    T acc = T(0);
    for (int64_t r = 0; r < A.rsize; r++) acc = fadd(acc, ld<LT>(A.in_dt, A.in, ib + r_offset(A, r)));
    ((T *)A.out)[ob] = acc;
}

__device__ inline int64_t group_off(const RedArgs &A, int64_t q) {
    if (A.nr == 0) return w_offset(A, q);
    const int64_t r = q / A.wsize, w = q - r * A.wsize;
    return w_offset(A, w) + r_offset(A, r);
}

template <int LT> __device__ void iadd_tile(const BatchCall &C, char *partials, int64_t g, uint64_t *red) {
    const RedArgs &A = C.red;
    const int64_t o = g / C.nseg, sg = g - o * C.nseg;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    const int64_t q0 = sg * kIntSeg, q1 = min(A.group, q0 + kIntSeg);
    uint64_t acc = 0;
    for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) acc += (uint64_t)ld<LT>(A.in_dt, A.in, ib + group_off(A, q));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        using ST = typename dt_type<LT>::T;
        if (C.direct) ((ST *)A.out)[ob] = (ST)red[0];
        else ((uint64_t *)(partials + C.part_off))[g] = red[0];
    }
}

template <int LT> __device__ void iadd_fold(const BatchCall &C, const char *partials) {
    using ST = typename dt_type<LT>::T;
    const RedArgs &A = C.red;
    for (int64_t o = threadIdx.x; o < A.nout; o += blockDim.x) {
        const uint64_t *p = (const uint64_t *)(partials + C.part_off) + o * C.nseg;
        uint64_t acc = 0;
        for (int64_t i = 0; i < C.nseg; i++) acc += p[i];
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((ST *)A.out)[ob] = (ST)acc;
    }
}

__device__ void fill_zero_tile(const BatchCall &C, int64_t t) {
    const RedArgs &A = C.red;
    const int64_t o = t * kNarrowOutputs + threadIdx.x;
    if (o >= A.nout) return;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    switch (itemsize(A.out_dt)) {
        case 1: ((uint8_t *)A.out)[ob] = 0; break;
        case 2: ((uint16_t *)A.out)[ob] = 0; break;
        case 4: ((uint32_t *)A.out)[ob] = 0; break;
        default: ((uint64_t *)A.out)[ob] = 0; break;
    }
}

__global__ __launch_bounds__(kBatchThreads) void batch_persistent(const BatchCall *calls, const int64_t *tile_begin,
                                                                   int32_t ncalls, int64_t ntiles, uint32_t *counters,
                                                                   char *partials) {
    __shared__ PwScratch S;
    __shared__ uint64_t ired[kBatchThreads];
    __shared__ int64_t sh_tile;
    __shared__ int32_t sh_call;
    for (;;) {
        if (threadIdx.x == 0) {
            const int64_t t = (int64_t)atomicAdd(&counters[0], 1u);
            int32_t c = -1;
            if (t < ntiles) {  // owning call: last c with tile_begin[c] <= t
                int32_t lo = 0, hi = ncalls - 1;
                while (lo < hi) {
                    const int32_t mid = (lo + hi + 1) >> 1;
                    if (tile_begin[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                c = lo;
            }
            sh_tile = t;
            sh_call = c;
        }
        __syncthreads();
        const int64_t t = sh_tile;
        const int32_t c = sh_call;
        if (c < 0) break;  // every tile has been handed out
        const BatchCall &C = calls[c];
        const int64_t lt = t - tile_begin[c];
        switch (C.kind) {
            case BK_ELEMENTWISE: ew_tile(C.ew, lt); break;
            case BK_FADD_WIDE:
                if (C.red.out_dt == IONP_FLOAT64) fadd_window<double>(C, partials, lt, S);
                else fadd_window<float>(C, partials, lt, S);
                break;
            case BK_FADD_NARROW:
                if (C.red.out_dt == IONP_FLOAT64) fadd_narrow<double>(C, lt);
                else fadd_narrow<float>(C, lt);
                break;
            case BK_IADD:
                switch (C.red.out_dt) {
                    case IONP_INT64: iadd_tile<IONP_INT64>(C, partials, lt, ired); break;
                    default: iadd_tile<IONP_UINT64>(C, partials, lt, ired); break;
                }
                break;
            default: fill_zero_tile(C, lt); break;
        }
        __syncthreads();
    }
}

// Second launch: calls whose outputs span several tiles fold their partials
// (float: acc = 0; acc = acc + partial in window order; integer: any order).
// Stream order makes every partial of the first launch visible.
__global__ __launch_bounds__(kBatchThreads) void batch_fold(const BatchCall *calls, const int32_t *fold_list,
                                                            int32_t nfold, const char *partials) {
    for (int32_t f = blockIdx.x; f < nfold; f += gridDim.x) {
        const BatchCall &C = calls[fold_list[f]];
        if (C.kind == BK_FADD_WIDE) {
            if (C.red.out_dt == IONP_FLOAT64) fadd_fold<double>(C, partials);
            else fadd_fold<float>(C, partials);
        } else if (C.red.out_dt == IONP_INT64) {
            iadd_fold<IONP_INT64>(C, partials);
        } else {
            iadd_fold<IONP_UINT64>(C, partials);
        }
    }
}

ionp_status plan_call(const ionp_call &c, BatchCall *B, int64_t *bytes) {
    std::memset(B, 0, sizeof *B);
    ionp_status s;
    if (c.kind == IONP_CALL_BINARY || c.kind == IONP_CALL_WHERE || c.kind == IONP_CALL_CAST) {
        ionp_array_desc t;
        int nin, op, loop;
        const ionp_array_desc *ins[3] = {&c.in[0], &c.in[1], &c.in[2]};
        if (c.kind == IONP_CALL_BINARY) {
            if ((s = ionp_plan_binary(c.op, &c.in[0], &c.in[1], &t))) return s;
            int odt;
            if ((s = binary_types(c.op, c.in[0].dtype, c.in[1].dtype, &loop, &odt))) return s;
            nin = 2;
            op = c.op;
        } else if (c.kind == IONP_CALL_WHERE) {
            if ((s = ionp_plan_where(&c.in[0], &c.in[1], &c.in[2], &t))) return s;
            nin = 3;
            op = OP_WHERE;
            loop = c.out.dtype;
        } else {
            if ((s = ionp_plan_cast(&c.in[0], c.out.dtype, &t))) return s;
            nin = 1;
            op = OP_CAST;
            loop = c.out.dtype;
        }
        if ((s = check_desc(&c.out, "out"))) return s;
        bool same = c.out.ndim == t.ndim;
        for (int i = 0; same && i < t.ndim; i++) same = c.out.shape[i] == t.shape[i];
        if (!same) return set_error(IONP_ERR_VALUE, "batch call output has the wrong shape");
        if ((s = plan_elementwise(op, nin, ins, &c.out, loop, &B->ew))) return s;
        B->ew.mixed_su = c.kind == IONP_CALL_BINARY && op >= IONP_LESS &&
                         ((c.in[0].dtype == IONP_INT64 && c.in[1].dtype == IONP_UINT64) ||
                          (c.in[0].dtype == IONP_UINT64 && c.in[1].dtype == IONP_INT64));
        if (B->ew.total >= (1ll << 31)) return set_error(IONP_ERR_UNSUPPORTED, "batch call too large");
        for (int d = 0; d < B->ew.ndim; d++) B->ew.div[d] = FastDiv((uint32_t)B->ew.shape[d]);
        B->kind = BK_ELEMENTWISE;
        B->tiles = (B->ew.total + kEwTile - 1) / kEwTile;
        *bytes = B->ew.total * itemsize(c.out.dtype);
        for (int k = 0; k < nin; k++) *bytes += desc_size(&c.in[k]) * itemsize(c.in[k].dtype);
        return IONP_OK;
    }
    if (c.kind != IONP_CALL_REDUCE) return set_error(IONP_ERR_VALUE, "unknown batch call kind");
    if (c.op != IONP_ADD) return set_error(IONP_ERR_UNSUPPORTED, "batched reductions support add.reduce");
    ionp_array_desc t;
    if ((s = ionp_plan_reduce(c.op, &c.in[0], c.axes, c.naxes, c.keepdims, c.out.dtype, &t))) return s;
    bool mask[IONP_MAX_DIMS];
    normalize_axes(c.axes, c.naxes, c.in[0].ndim, mask);
    ReducePlan p;
    if ((s = plan_reduce(c.op, &c.in[0], mask, &c.out, &p))) return s;
    RedArgs &A = B->red;
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
    A.bufsize = ionp_get_bufsize();
    A.nwin = p.wsize > 0 ? (p.wsize + A.bufsize - 1) / A.bufsize : 0;
    A.in = p.in;
    A.out = p.out;
    for (int d = 0; d < p.nk; d++) {
        A.k_shape[d] = p.k_shape[d];
        A.k_in_st[d] = p.k_in_st[d];
        A.k_out_st[d] = p.k_out_st[d];
        A.k_div[d] = FastDiv((uint32_t)p.k_shape[d]);
    }
    for (int d = 0; d < p.nw; d++) {
        A.w_shape[d] = p.w_shape[d];
        A.w_st[d] = p.w_st[d];
        A.w_div[d] = FastDiv((uint32_t)p.w_shape[d]);
    }
    for (int d = 0; d < p.nr; d++) {
        A.r_shape[d] = p.r_shape[d];
        A.r_st[d] = p.r_st[d];
        A.r_div[d] = FastDiv((uint32_t)p.r_shape[d]);
    }
    if (p.nout >= (1ll << 31) || p.group >= (1ll << 31)) return set_error(IONP_ERR_UNSUPPORTED, "batch call too large");
    *bytes = desc_size(&c.in[0]) * itemsize(c.in[0].dtype) + p.nout * itemsize(p.out_dt);
    if (p.nout == 0) {
        B->kind = BK_FILL_ZERO;
        B->tiles = 0;
    } else if (p.group == 0) {
        B->kind = BK_FILL_ZERO;
        B->tiles = (p.nout + kNarrowOutputs - 1) / kNarrowOutputs;
    } else if (is_float_dt(p.out_dt)) {
        if (p.nw == 0) {
            B->kind = BK_FADD_NARROW;
            B->tiles = (p.nout + kNarrowOutputs - 1) / kNarrowOutputs;
        } else {
            B->kind = BK_FADD_WIDE;
            B->tiles = p.nout * p.rsize * A.nwin;
            B->direct = p.rsize == 1 && A.nwin == 1;
        }
    } else {
        B->kind = BK_IADD;
        B->nseg = (p.group + kIntSeg - 1) / kIntSeg;
        B->tiles = p.nout * B->nseg;
        B->direct = B->nseg == 1;
    }
    return IONP_OK;
}

static ionp_status run_one(const ionp_call &c, hipStream_t s, const ionp_config *cfg = nullptr) {
    switch (c.kind) {
        case IONP_CALL_BINARY: return ionp_binary(c.op, &c.in[0], &c.in[1], &c.out, cfg, s);
        case IONP_CALL_WHERE: return ionp_where(&c.in[0], &c.in[1], &c.in[2], &c.out, cfg, s);
        case IONP_CALL_CAST: return ionp_cast(&c.in[0], &c.out, cfg, s);
        default: return ionp_reduce(c.op, &c.in[0], c.axes, c.naxes, c.keepdims, &c.out, cfg, s);
    }
}

}  // namespace ionp

using namespace ionp;

extern "C" {

ionp_status ionp_batch_create(const ionp_call *calls, int32_t n, void *stream, ionp_batch **out) {
    if (!out || (n > 0 && !calls) || n < 0) return set_error(IONP_ERR_VALUE, "bad batch arguments");
    *out = nullptr;
    std::vector<BatchCall> table(n > 0 ? n : 1);
    std::vector<int64_t> begin(n + 1, 0);
    std::vector<int32_t> fold_list;
    int64_t bytes = 0, part = 0;
    size_t scratch_need = 0;
    for (int32_t i = 0; i < n; i++) {
        int64_t b = 0;
        ionp_status s = plan_call(calls[i], &table[i], &b);
        if (s) {
            return set_error(s, "batch call " + std::to_string(i) + ": " + last_error());
        }
        BatchCall &B = table[i];
        if ((B.kind == BK_FADD_WIDE || B.kind == BK_IADD) && !B.direct) {
            B.part_off = part;
            part += ((B.tiles * 8) + 255) / 256 * 256;
            scratch_need = std::max(scratch_need, (size_t)(B.tiles * 16));
            fold_list.push_back(i);
        }
        begin[i + 1] = begin[i] + B.tiles;
        bytes += b;
    }
    ionp_batch *bt = new ionp_batch();
    bt->n = n;
    bt->ntiles = begin[n];
    bt->bytes = bytes;
    bt->part_bytes = part;
    bt->calls.assign(calls, calls + n);
    bt->graph_ready = false;
    bt->scratch_need = scratch_need;
    hipStream_t s = (hipStream_t)stream;
    bt->nfold = (int32_t)fold_list.size();
    bool ok = hipMalloc(&bt->d_calls, sizeof(BatchCall) * table.size()) == hipSuccess &&
              hipMalloc(&bt->d_tile_begin, sizeof(int64_t) * (n + 1)) == hipSuccess &&
              hipMalloc(&bt->d_counters, sizeof(uint32_t) * 64) == hipSuccess &&
              hipMalloc(&bt->d_fold_list, sizeof(int32_t) * (fold_list.size() + 1)) == hipSuccess &&
              hipMalloc(&bt->d_partials, part > 0 ? part : 256) == hipSuccess;
    ok = ok && hipMemcpyAsync(bt->d_calls, table.data(), sizeof(BatchCall) * table.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(bt->d_tile_begin, begin.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s) == hipSuccess &&
         (fold_list.empty() || hipMemcpyAsync(bt->d_fold_list, fold_list.data(), sizeof(int32_t) * fold_list.size(),
                                              hipMemcpyHostToDevice, s) == hipSuccess) &&
         hipStreamSynchronize(s) == hipSuccess;
    if (!ok) {
        ionp_batch_destroy(bt);
        return set_error(IONP_ERR_DEVICE, "batch allocation failed");
    }
    *out = bt;
    return IONP_OK;
}

ionp_status ionp_batch_run(ionp_batch *bt, int32_t mode, void *stream) {
    if (!bt) return set_error(IONP_ERR_VALUE, "null batch");
    hipStream_t s = (hipStream_t)stream;
    if (mode == IONP_BATCH_PER_CALL) {
        for (const ionp_call &c : bt->calls) {
            ionp_status st = run_one(c, s);
            if (st) return st;
        }
        return IONP_OK;
    }
    if (mode == IONP_BATCH_GRAPH) {
        if (!bt->graph_ready) {
            // one eager pass first: it sizes the library scratch, so nothing
            // allocates inside the capture
            for (const ionp_call &c : bt->calls) {
                ionp_status st0 = run_one(c, s);
                if (st0) return st0;
            }
            (void)num_cus();
            // capture on a private stream (the legacy default stream can not be
            // captured); the graph is then launched on the caller's stream
            hipStream_t cs;
            if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
                return set_error(IONP_ERR_DEVICE, "hipStreamCreate failed");
            if (hipStreamSynchronize(s) != hipSuccess || hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                (void)hipStreamDestroy(cs);
                return set_error(IONP_ERR_DEVICE, "hipStreamBeginCapture failed");
            }
            ionp_status st = IONP_OK;
            for (const ionp_call &c : bt->calls)
                if (!st) st = run_one(c, cs);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(cs, &g);
            (void)hipStreamDestroy(cs);
            if (st) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            if (e != hipSuccess) return set_error(IONP_ERR_DEVICE, "hipStreamEndCapture failed");
            if (hipGraphInstantiate(&bt->graph, g, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(g);
                return set_error(IONP_ERR_DEVICE, "hipGraphInstantiate failed");
            }
            (void)hipGraphDestroy(g);
            bt->graph_ready = true;
        }
        return hipGraphLaunch(bt->graph, s) == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "hipGraphLaunch failed");
    }
    if (mode != IONP_BATCH_PERSISTENT) return set_error(IONP_ERR_VALUE, "unknown batch mode");
    if (bt->n == 0 || bt->ntiles == 0) return IONP_OK;
    hipLaunchKernelGGL(batch_zero, dim3(1), dim3(64), 0, s, bt->d_counters, (int64_t)64);
    // persistent grid: a few workgroups per CU, never more than the tiles
    int64_t grid = (int64_t)num_cus() * 4;
    if (grid > bt->ntiles) grid = bt->ntiles;
    hipLaunchKernelGGL(batch_persistent, dim3((unsigned)grid), dim3(kBatchThreads), 0, s, bt->d_calls, bt->d_tile_begin,
                       bt->n, bt->ntiles, bt->d_counters, bt->d_partials);
    if (bt->nfold)
        hipLaunchKernelGGL(batch_fold, dim3((unsigned)std::min<int32_t>(bt->nfold, num_cus() * 4)), dim3(kBatchThreads), 0, s,
                           bt->d_calls, bt->d_fold_list, bt->nfold, bt->d_partials);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, hipGetErrorString(e));
}

ionp_status ionp_batch_info(const ionp_batch *bt, int64_t *ntiles, int64_t *bytes) {
    if (!bt) return set_error(IONP_ERR_VALUE, "null batch");
    if (ntiles) *ntiles = bt->ntiles;
    if (bytes) *bytes = bt->bytes;
    return IONP_OK;
}

ionp_status ionp_autotune(const ionp_call *calls, int32_t n, const ionp_config *cfgs, int32_t ncfg, int32_t reps,
                          float *times_us, void *stream) {
    if (n < 0 || ncfg <= 0 || reps <= 0 || (n > 0 && (!calls || !cfgs || !times_us)))
        return set_error(IONP_ERR_VALUE, "bad autotune arguments");
    hipStream_t s = (hipStream_t)stream;
    const int nev = ncfg * reps;
    std::vector<hipEvent_t> ev(2 * (size_t)nev);
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) return set_error(IONP_ERR_DEVICE, "hipEventCreate failed");
    ionp_status st = IONP_OK;
    for (int32_t i = 0; i < n && !st; i++) {
        // warm the call (first-touch, scratch growth) with the default configuration
        if ((st = run_one(calls[i], s))) break;
        std::vector<char> usable(ncfg, 1);
        for (int32_t j = 0; j < ncfg; j++)
            for (int32_t r = 0; r < reps; r++) {
                (void)hipEventRecord(ev[2 * (j * reps + r)], s);
                if (run_one(calls[i], s, &cfgs[j]) != IONP_OK) usable[j] = 0;
                (void)hipEventRecord(ev[2 * (j * reps + r) + 1], s);
            }
        if (hipStreamSynchronize(s) != hipSuccess) {
            st = set_error(IONP_ERR_DEVICE, "autotune run failed");
            break;
        }
        for (int32_t j = 0; j < ncfg; j++) {
            float best = -1.0f;
            for (int32_t r = 0; r < reps && usable[j]; r++) {
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, ev[2 * (j * reps + r)], ev[2 * (j * reps + r) + 1]);
                if (best < 0.0f || ms * 1000.0f < best) best = ms * 1000.0f;
            }
            times_us[(size_t)i * ncfg + j] = best;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return st;
}

void ionp_batch_destroy(ionp_batch *bt) {
    if (!bt) return;
    if (bt->graph_ready) (void)hipGraphExecDestroy(bt->graph);
    if (bt->d_calls) (void)hipFree(bt->d_calls);
    if (bt->d_tile_begin) (void)hipFree(bt->d_tile_begin);
    if (bt->d_counters) (void)hipFree(bt->d_counters);
    if (bt->d_fold_list) (void)hipFree(bt->d_fold_list);
    if (bt->d_partials) (void)hipFree(bt->d_partials);
    delete bt;
}

}  // extern "C"
