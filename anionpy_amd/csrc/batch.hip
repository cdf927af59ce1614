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
#include <algorithm>
#include <type_traits>
#include <vector>

#include "device_ops.h"

namespace ionp {

enum : int32_t {
    BK_ELEMENTWISE = 0,  // binary / where / cast
    BK_FADD_WIDE = 1,    // float add.reduce, windows over W
    BK_FADD_NARROW = 2,  // float add.reduce, W empty
    BK_IADD = 3,         // integer / bool add.reduce (order-free), segments of a group
    BK_FILL_ZERO = 4,    // empty reduction groups
    BK_FADD_SMALL = 5,   // float add.reduce, windows of <= 128: one thread per output
    BK_IADD_SMALL = 6,   // integer / bool add.reduce: one thread per output
    BK_FADD_NARROW_WAVE = 7,  // float add.reduce, W empty, long chains: one wavefront per output
    BK_IADD_WAVE = 8,    // integer / bool add.reduce, groups of 65 .. kIntWaveMax: one wavefront per output
    BK_SIDE = 9,         // float16 call: its own single-call launches, ahead of the persistent kernel
    BK_EW_FLAT = 10,     // binary / where / cast over one contiguous, aligned dim: nontemporal vectors
    BK_EW_SIZED = 11,    // binary / where, inputs of 4- / 8-byte items, 31-bit offsets (sized_ok)
};

constexpr int kBatchThreads = 256;
constexpr int64_t kEwTileDefault = kBatchThreads * 64;  // elements per elementwise tile (ionp_batch_config)
constexpr int64_t kIntSegDefault = 8192;                // elements per integer-reduce tile (ionp_batch_config)
constexpr int64_t kNarrowOutputs = kBatchThreads;       // outputs per narrow tile (a thread each)
constexpr int32_t kClaimDefault = 4;                     // tiles per claim (ionp_batch_config)
constexpr int64_t kSerialChain = 2048;                   // a tile with a longer dependent chain is claimed alone
constexpr int64_t kIntWaveMax = 16384;           // longest group a wavefront reduces alone

struct BatchCall {
    int32_t kind;
    int32_t direct;    // reduce result written by the tile itself (one tile per output)
    int64_t tiles;     // tiles of this call
    int64_t tile;      // elements per tile (elementwise) or per segment (integer reduce)
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
    int32_t *d_tile_map;        // device: per tile, (owning call, tile index within the call)
    uint32_t *d_counters;       // device: [0] tile counter
    int32_t *d_fold_list;       // device: calls whose partials need a fold
    int32_t nfold;
    char *d_partials;           // device: reduction partials
    std::vector<ionp_call> calls;  // host copy for GRAPH / PER_CALL modes
    hipGraphExec_t graph;
    bool graph_ready;
    size_t scratch_need;
    // GRAPH mode's own scratch (partials, converted inputs, fold slots): the
    // captured kernels point into it, so no later call can free or refill it
    char *d_graph_scratch;
    ionp::ScratchArena graph_arena;
    std::vector<int32_t> side;  // float16 calls (original order): single-call launches ahead of the persistent kernel
    int32_t claim;              // tiles per claim of the persistent kernel
    int64_t nserial;            // leading tiles with long serial chains (claimed one at a time)
};

namespace ionp {

__global__ void batch_zero(uint32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// K loads of one operand (dtype dt, converted to the loop dtype LT), issued
// back to back: the dtype switch sits outside the loads.
template <int LT, int K>
__device__ inline void load_k(int dt, const char *p, const int64_t (&off)[K], typename dt_type<LT>::T (&v)[K]) {
    in_type_dispatch(dt, [&](auto it) {
        using IT = decltype(it);
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = ld_t<LT, IT>(p, off[k]);
        return 0;
    });
}

template <typename T, int K, typename XT>
__device__ inline void store_k(int dt, char *p, const XT (&off)[K], const T (&v)[K], const bool (&ok)[K]) {
    auto st = [&](auto ot, auto cv) {
        using OT = decltype(ot);
#pragma unroll
        for (int k = 0; k < K; k++)
            if (ok[k]) ((OT *)p)[off[k]] = cv(v[k]);
        return 0;
    };
    switch (dt) {
        case IONP_BOOL: st(uint8_t(), [](T x) { return cvt<IONP_BOOL>(x); }); break;
        case IONP_INT8: st(int8_t(), [](T x) { return cvt<IONP_INT8>(x); }); break;
        case IONP_UINT8: st(uint8_t(), [](T x) { return cvt<IONP_UINT8>(x); }); break;
        case IONP_INT16: st(int16_t(), [](T x) { return cvt<IONP_INT16>(x); }); break;
        case IONP_UINT16: st(uint16_t(), [](T x) { return cvt<IONP_UINT16>(x); }); break;
        case IONP_INT32: st(int32_t(), [](T x) { return cvt<IONP_INT32>(x); }); break;
        case IONP_UINT32: st(uint32_t(), [](T x) { return cvt<IONP_UINT32>(x); }); break;
        case IONP_INT64: st(int64_t(), [](T x) { return cvt<IONP_INT64>(x); }); break;
        case IONP_UINT64: st(uint64_t(), [](T x) { return cvt<IONP_UINT64>(x); }); break;
        case IONP_FLOAT32: st(float(), [](T x) { return cvt<IONP_FLOAT32>(x); }); break;
        default: st(double(), [](T x) { return cvt<IONP_FLOAT64>(x); }); break;
    }
}

// Elements [e0, e1) of an elementwise call.  Thread t takes elements
// e0 + t + j * blockDim.x; K of them at a time have their offsets computed,
// their loads issued together, then their results stored (K loads per
// operand in flight instead of one).  A slot past e1 re-reads element e0 and
// stores nothing.
//
// ND > 0: the call has exactly ND dims, so the index loop unrolls and the
// call's shape / stride fields stay in scalar registers for the whole tile.
constexpr int kRangeK = 4;  // elements per thread in flight in the generic tiles

// tid / nthreads: the threads sharing the range (default: the workgroup;
// one thread alone with 0 / 1)
template <int LT, int ND>
__device__ void ew_range(const EwArgs &A, uint32_t e0, uint32_t e1, uint32_t tid = threadIdx.x,
                         uint32_t nthreads = blockDim.x) {
    using T = typename dt_type<LT>::T;
    constexpr int K = kRangeK;
    const int nin = A.nin;
    const int nd = ND > 0 ? ND : A.ndim;
    for (uint32_t base = e0 + tid; base < e1; base += K * nthreads) {
        int64_t off[3][K], oo[K];
        bool ok[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            uint32_t e = base + k * nthreads;
            ok[k] = e < e1;
            uint32_t rem = ok[k] ? e : e0;
            int64_t o0 = 0, o1 = 0, o2 = 0, ou = 0;
#pragma unroll
            for (int d = 0; d < (ND > 0 ? ND : kMaxKDims); d++) {
                if (ND == 0 && d >= nd) break;
                uint32_t c;
                if (d + 1 < nd) {
                    const uint32_t q = A.div[d].div(rem);
                    c = rem - q * (uint32_t)A.shape[d];
                    rem = q;
                } else {
                    c = rem;
                }
                o0 += (int64_t)c * A.in_st[0][d];
                if (nin > 1) o1 += (int64_t)c * A.in_st[1][d];
                if (nin > 2) o2 += (int64_t)c * A.in_st[2][d];
                ou += (int64_t)c * A.out_st[d];
            }
            off[0][k] = o0;
            off[1][k] = o1;
            off[2][k] = o2;
            oo[k] = ou;
        }
        T r[K];
        if (A.op == OP_CAST) {
            load_k<LT, K>(A.in_dt[0], A.in[0], off[0], r);
        } else if (A.op == OP_WHERE) {
            uint8_t c[K];
            T x[K], y[K];
            load_k<IONP_BOOL, K>(A.in_dt[0], A.in[0], off[0], c);
            load_k<LT, K>(A.in_dt[1], A.in[1], off[1], x);
            load_k<LT, K>(A.in_dt[2], A.in[2], off[2], y);
#pragma unroll
            for (int k = 0; k < K; k++) r[k] = c[k] ? x[k] : y[k];
        } else if (A.op >= IONP_LESS) {
            uint8_t cr[K];
            if (A.mixed_su) {
                const bool first = A.in_dt[0] == IONP_INT64;
                int64_t sv[K], so[K], uo[K];
                uint64_t uv[K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    so[k] = first ? off[0][k] : off[1][k];
                    uo[k] = first ? off[1][k] : off[0][k];
                }
                load_k<IONP_INT64, K>(IONP_INT64, A.in[first ? 0 : 1], so, sv);
                load_k<IONP_UINT64, K>(IONP_UINT64, A.in[first ? 1 : 0], uo, uv);
                const int op = first ? A.op : swap_cmp(A.op);
#pragma unroll
                for (int k = 0; k < K; k++) cr[k] = (uint8_t)cmp_su(op, sv[k], uv[k]);
            } else {
                T a[K], b[K];
                load_k<LT, K>(A.in_dt[0], A.in[0], off[0], a);
                load_k<LT, K>(A.in_dt[1], A.in[1], off[1], b);
#pragma unroll
                for (int k = 0; k < K; k++) cr[k] = (uint8_t)cmp_op<T>(A.op, a[k], b[k]);
            }
            store_k<uint8_t, K>(A.out_dt, A.out, oo, cr, ok);
            continue;
        } else {
            T a[K], b[K];
            load_k<LT, K>(A.in_dt[0], A.in[0], off[0], a);
            load_k<LT, K>(A.in_dt[1], A.in[1], off[1], b);
#pragma unroll
            for (int k = 0; k < K; k++) r[k] = arith_op<LT>(A.op, a[k], b[k]);
        }
        store_k<T, K>(A.out_dt, A.out, oo, r, ok);
    }
}

template <int LT> __device__ inline void ew_range_nd(const EwArgs &A, uint32_t e0, uint32_t e1) {
    switch (A.ndim) {
        case 1: ew_range<LT, 1>(A, e0, e1); break;
        case 2: ew_range<LT, 2>(A, e0, e1); break;
        case 3: ew_range<LT, 3>(A, e0, e1); break;
        case 4: ew_range<LT, 4>(A, e0, e1); break;
        default: ew_range<LT, 0>(A, e0, e1); break;
    }
}

__device__ void ew_tile(const EwArgs &A, int64_t t, int64_t tile) {
    const uint32_t e0 = (uint32_t)(t * tile), e1 = (uint32_t)min(A.total, t * tile + tile);
    switch (A.loop_dt) {  // the common loop dtypes get rank-specialised index code
        case IONP_FLOAT64: ew_range_nd<IONP_FLOAT64>(A, e0, e1); break;
        case IONP_FLOAT32: ew_range_nd<IONP_FLOAT32>(A, e0, e1); break;
        case IONP_INT64: ew_range_nd<IONP_INT64>(A, e0, e1); break;
        case IONP_BOOL: ew_range<IONP_BOOL, 0>(A, e0, e1); break;
        case IONP_INT8: ew_range<IONP_INT8, 0>(A, e0, e1); break;
        case IONP_UINT8: ew_range<IONP_UINT8, 0>(A, e0, e1); break;
        case IONP_INT16: ew_range<IONP_INT16, 0>(A, e0, e1); break;
        case IONP_UINT16: ew_range<IONP_UINT16, 0>(A, e0, e1); break;
        case IONP_INT32: ew_range<IONP_INT32, 0>(A, e0, e1); break;
        case IONP_UINT32: ew_range<IONP_UINT32, 0>(A, e0, e1); break;
        default: ew_range<IONP_UINT64, 0>(A, e0, e1); break;
    }
}

// ---- the sized tile path ----------------------------------------------------
// For calls whose inputs have 4- or 8-byte items (where's condition: bool)
// and whose offsets fit in 31 bits (sized_ok).  The item sizes are template
// arguments, so each operand's K loads are plain fixed-width loads issued
// back to back for every operand, and the dtype conversion runs after all of
// them.  (In ew_range each operand's loads sit in a dtype switch with their
// conversion, which makes the next operand's loads wait for that data.)
template <int S> struct uint_sz;
template <> struct uint_sz<1> { using T = uint8_t; };
template <> struct uint_sz<4> { using T = uint32_t; };
template <> struct uint_sz<8> { using T = uint64_t; };

// K raw items (unsigned U of the dtype's size) of dtype dt converted to the
// loop dtype LT as ld_t converts them.  dt's item size must be sizeof(U)
// (bool: U is uint8_t).
template <int LT, int K, typename U>
__device__ inline void cvt_raw_k(int dt, const U (&r)[K], typename dt_type<LT>::T (&v)[K]) {
    in_type_dispatch(dt, [&](auto it) {
        using IT = decltype(it);
        if constexpr (std::is_same<IT, bool>::value) {
#pragma unroll
            for (int k = 0; k < K; k++) v[k] = cvt<LT>((uint8_t)(r[k] != 0));
        } else if constexpr (sizeof(IT) == sizeof(U)) {
#pragma unroll
            for (int k = 0; k < K; k++) v[k] = cvt<LT>(__builtin_bit_cast(IT, r[k]));
        } else {
            __builtin_unreachable();  // sized_ok admits only dtypes of U's size
        }
        return 0;
    });
}

// Elements [e0, e1) as ew_range, for NIN inputs of item sizes SA, SB, SC
// over exactly ND dims (the host pads the call's dims to 2 or 4).
template <int LT, int NIN, int SA, int SB, int SC, int ND>
__device__ void ew_range_sz(const EwArgs &A, uint32_t e0, uint32_t e1) {
    using T = typename dt_type<LT>::T;
    using UA = typename uint_sz<SA>::T;
    using UB = typename uint_sz<(NIN > 1 ? SB : 1)>::T;
    using UC = typename uint_sz<(NIN > 2 ? SC : 1)>::T;
    constexpr int K = kRangeK;
    for (uint32_t base = e0 + threadIdx.x; base < e1; base += K * blockDim.x) {
        int32_t off[3][K], oo[K];
        bool ok[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            uint32_t e = base + k * blockDim.x;
            ok[k] = e < e1;
            uint32_t rem = ok[k] ? e : e0;
            int32_t o0 = 0, o1 = 0, o2 = 0, ou = 0;
#pragma unroll
            for (int d = 0; d < ND; d++) {
                uint32_t c;
                if (d + 1 < ND) {
                    const uint32_t q = A.div[d].div(rem);
                    c = rem - q * (uint32_t)A.shape[d];
                    rem = q;
                } else {
                    c = rem;
                }
                o0 += (int32_t)c * (int32_t)A.in_st[0][d];
                if constexpr (NIN > 1) o1 += (int32_t)c * (int32_t)A.in_st[1][d];
                if constexpr (NIN > 2) o2 += (int32_t)c * (int32_t)A.in_st[2][d];
                ou += (int32_t)c * (int32_t)A.out_st[d];
            }
            off[0][k] = o0;
            off[1][k] = o1;
            off[2][k] = o2;
            oo[k] = ou;
        }
        UA ra[K];
        UB rb[K];
        UC rc[K];
#pragma unroll
        for (int k = 0; k < K; k++) ra[k] = ((const UA *)A.in[0])[off[0][k]];
        if constexpr (NIN > 1) {
#pragma unroll
            for (int k = 0; k < K; k++) rb[k] = ((const UB *)A.in[1])[off[1][k]];
        }
        if constexpr (NIN > 2) {
#pragma unroll
            for (int k = 0; k < K; k++) rc[k] = ((const UC *)A.in[2])[off[2][k]];
        }
        T r[K];
        if constexpr (NIN == 1) {
            cvt_raw_k<LT, K>(A.in_dt[0], ra, r);
        } else if constexpr (NIN == 3) {
            T x[K], y[K];
            cvt_raw_k<LT, K>(A.in_dt[1], rb, x);
            cvt_raw_k<LT, K>(A.in_dt[2], rc, y);
#pragma unroll
            for (int k = 0; k < K; k++) r[k] = ra[k] ? x[k] : y[k];
        } else {
            T a[K], b[K];
            cvt_raw_k<LT, K>(A.in_dt[0], ra, a);
            cvt_raw_k<LT, K>(A.in_dt[1], rb, b);
            if (A.op >= IONP_LESS) {
                uint8_t cr[K];
#pragma unroll
                for (int k = 0; k < K; k++) cr[k] = (uint8_t)cmp_op<T>(A.op, a[k], b[k]);
                store_k<uint8_t, K>(A.out_dt, A.out, oo, cr, ok);
                continue;
            }
#pragma unroll
            for (int k = 0; k < K; k++) r[k] = arith_op<LT>(A.op, a[k], b[k]);
        }
        store_k<T, K>(A.out_dt, A.out, oo, r, ok);
    }
}

template <int LT, int ND> __device__ inline void ew_sized(const EwArgs &A, uint32_t e0, uint32_t e1) {
    const bool a8 = itemsize(A.in_dt[0]) == 8, b8 = itemsize(A.in_dt[1]) == 8,
               c8 = A.nin > 2 && itemsize(A.in_dt[2]) == 8;
    if (A.nin == 2) {
        if (a8 && b8) ew_range_sz<LT, 2, 8, 8, 0, ND>(A, e0, e1);
        else if (a8) ew_range_sz<LT, 2, 8, 4, 0, ND>(A, e0, e1);
        else if (b8) ew_range_sz<LT, 2, 4, 8, 0, ND>(A, e0, e1);
        else ew_range_sz<LT, 2, 4, 4, 0, ND>(A, e0, e1);
    } else {
        if (b8 && c8) ew_range_sz<LT, 3, 1, 8, 8, ND>(A, e0, e1);
        else if (b8) ew_range_sz<LT, 3, 1, 8, 4, ND>(A, e0, e1);
        else if (c8) ew_range_sz<LT, 3, 1, 4, 8, ND>(A, e0, e1);
        else ew_range_sz<LT, 3, 1, 4, 4, ND>(A, e0, e1);
    }
}

template <int LT> __device__ inline void ew_sized_nd(const EwArgs &A, uint32_t e0, uint32_t e1) {
    if (A.ndim == 2) ew_sized<LT, 2>(A, e0, e1);
    else ew_sized<LT, 4>(A, e0, e1);
}

__device__ void ew_sized_tile(const EwArgs &A, int64_t t, int64_t tile) {
    const uint32_t e0 = (uint32_t)(t * tile), e1 = (uint32_t)min(A.total, t * tile + tile);
    switch (A.loop_dt) {
        case IONP_FLOAT64: ew_sized_nd<IONP_FLOAT64>(A, e0, e1); break;
        case IONP_FLOAT32: ew_sized_nd<IONP_FLOAT32>(A, e0, e1); break;
        default: ew_sized_nd<IONP_INT64>(A, e0, e1); break;
    }
}

// ---------------------------------------------------------------------------
// Vector elementwise tiles: calls whose innermost coalesced dim has stride 1
// or 0 in every input and stride 1 in the output (any outer dims).  A
// thread takes 4 consecutive elements of that dim per vector: an operand of
// stride 1 is read with one 4-element load in its own dtype (declared with
// its element's alignment, so any row offset is valid), one of stride 0
// with a single load; the values are converted after the loads.  Two
// vectors are in flight per thread.  A vector that crosses a row or the
// tile's end takes the element-by-element path.
// ---------------------------------------------------------------------------
constexpr int kVecV = 4;

// v[0..3] = operand elements at off .. off + 3 (st 1) or off (st 0), as
// dtype LT.  AL: the address is aligned to the 4-element vector (a
// nontemporal load); otherwise to one element.
template <int LT, bool AL> __device__ inline void vload(int dt, const char *base, int64_t off, int64_t st,
                                                        typename dt_type<LT>::T (&v)[kVecV]) {
    in_type_dispatch(dt, [&](auto it) {
        using IT0 = decltype(it);
        constexpr bool B = std::is_same<IT0, bool>::value;
        using IT = typename std::conditional<B, uint8_t, IT0>::type;
        typedef IT VA __attribute__((ext_vector_type(kVecV)));
        typedef IT VU __attribute__((ext_vector_type(kVecV), aligned(sizeof(IT))));
        const IT *p = (const IT *)base + off;
        if (st == 0) {
            const IT x = *p;
#pragma unroll
            for (int e = 0; e < kVecV; e++) v[e] = B ? cvt<LT>((uint8_t)(x != 0)) : cvt<LT>(x);
        } else {
            VA x;
            if constexpr (AL) x = __builtin_nontemporal_load((const VA *)p);
            else x = *(const VU *)p;
#pragma unroll
            for (int e = 0; e < kVecV; e++) v[e] = B ? cvt<LT>((uint8_t)(x[e] != 0)) : cvt<LT>(x[e]);
        }
        return 0;
    });
}

template <typename T, bool AL> __device__ inline void vstore(int dt, char *base, int64_t off, const T (&v)[kVecV]) {
    auto st = [&](auto ot, auto cv) {
        using OT = decltype(ot);
        typedef OT VA __attribute__((ext_vector_type(kVecV)));
        typedef OT VU __attribute__((ext_vector_type(kVecV), aligned(sizeof(OT))));
        VA r;
#pragma unroll
        for (int e = 0; e < kVecV; e++) r[e] = cv(v[e]);
        if constexpr (AL) __builtin_nontemporal_store(r, (VA *)((OT *)base + off));
        else *(VU *)((OT *)base + off) = r;
        return 0;
    };
    switch (dt) {
        case IONP_BOOL: st(uint8_t(), [](T x) { return cvt<IONP_BOOL>(x); }); break;
        case IONP_INT8: st(int8_t(), [](T x) { return cvt<IONP_INT8>(x); }); break;
        case IONP_UINT8: st(uint8_t(), [](T x) { return cvt<IONP_UINT8>(x); }); break;
        case IONP_INT16: st(int16_t(), [](T x) { return cvt<IONP_INT16>(x); }); break;
        case IONP_UINT16: st(uint16_t(), [](T x) { return cvt<IONP_UINT16>(x); }); break;
        case IONP_INT32: st(int32_t(), [](T x) { return cvt<IONP_INT32>(x); }); break;
        case IONP_UINT32: st(uint32_t(), [](T x) { return cvt<IONP_UINT32>(x); }); break;
        case IONP_INT64: st(int64_t(), [](T x) { return cvt<IONP_INT64>(x); }); break;
        case IONP_UINT64: st(uint64_t(), [](T x) { return cvt<IONP_UINT64>(x); }); break;
        case IONP_FLOAT32: st(float(), [](T x) { return cvt<IONP_FLOAT32>(x); }); break;
        default: st(double(), [](T x) { return cvt<IONP_FLOAT64>(x); }); break;
    }
}

// Flat calls (one coalesced dim, every operand contiguous and aligned to 4
// of its elements): nontemporal 4-element vectors, two in flight per
// thread; the tile's last partial vector element by element.
template <int LT> __device__ void ew_flat(const EwArgs &A, uint32_t e0, uint32_t e1) {
    using T = typename dt_type<LT>::T;
    constexpr int K = 2, V = kVecV;
    const uint32_t step = blockDim.x * V;
    for (uint32_t base = e0 + threadIdx.x * V; base < e1; base += K * step) {
        if (base + (K - 1) * step + V <= e1) {
            if (A.op == OP_CAST) {
                T x[K][V];
#pragma unroll
                for (int q = 0; q < K; q++) vload<LT, true>(A.in_dt[0], A.in[0], base + q * step, 1, x[q]);
#pragma unroll
                for (int q = 0; q < K; q++) vstore<T, true>(A.out_dt, A.out, base + q * step, x[q]);
            } else if (A.op == OP_WHERE) {
                uint8_t c[K][V];
                T x[K][V], y[K][V];
#pragma unroll
                for (int q = 0; q < K; q++) {
                    vload<IONP_BOOL, true>(A.in_dt[0], A.in[0], base + q * step, 1, c[q]);
                    vload<LT, true>(A.in_dt[1], A.in[1], base + q * step, 1, x[q]);
                    vload<LT, true>(A.in_dt[2], A.in[2], base + q * step, 1, y[q]);
                }
#pragma unroll
                for (int q = 0; q < K; q++) {
                    T r[V];
#pragma unroll
                    for (int e = 0; e < V; e++) r[e] = c[q][e] ? x[q][e] : y[q][e];
                    vstore<T, true>(A.out_dt, A.out, base + q * step, r);
                }
            } else {
                T a[K][V], b[K][V];
#pragma unroll
                for (int q = 0; q < K; q++) {
                    vload<LT, true>(A.in_dt[0], A.in[0], base + q * step, 1, a[q]);
                    vload<LT, true>(A.in_dt[1], A.in[1], base + q * step, 1, b[q]);
                }
#pragma unroll
                for (int q = 0; q < K; q++) {
                    if (A.op >= IONP_LESS) {
                        uint8_t r[V];
#pragma unroll
                        for (int e = 0; e < V; e++) r[e] = (uint8_t)cmp_op<T>(A.op, a[q][e], b[q][e]);
                        vstore<uint8_t, true>(A.out_dt, A.out, base + q * step, r);
                    } else {
                        T r[V];
#pragma unroll
                        for (int e = 0; e < V; e++) r[e] = arith_op<LT>(A.op, a[q][e], b[q][e]);
                        vstore<T, true>(A.out_dt, A.out, base + q * step, r);
                    }
                }
            }
        } else {
            for (int q = 0; q < K; q++)
                for (uint32_t e = base + q * step; e < base + q * step + V && e < e1; e++) {
                    const int64_t isz0 = itemsize(A.in_dt[0]), isz1 = A.nin > 1 ? itemsize(A.in_dt[1]) : 0,
                                  isz2 = A.nin > 2 ? itemsize(A.in_dt[2]) : 0;
                    ew_element<LT>(A, A.in[0] + e * isz0, A.in[1] + e * isz1, A.in[2] + e * isz2,
                                   A.out + e * itemsize(A.out_dt));
                }
        }
    }
}

__device__ void ew_flat_tile(const EwArgs &A, int64_t t, int64_t tile) {
    const uint32_t e0 = (uint32_t)(t * tile), e1 = (uint32_t)min(A.total, t * tile + tile);
    switch (A.loop_dt) {
        case IONP_FLOAT64: ew_flat<IONP_FLOAT64>(A, e0, e1); break;
        case IONP_FLOAT32: ew_flat<IONP_FLOAT32>(A, e0, e1); break;
        case IONP_INT64: ew_flat<IONP_INT64>(A, e0, e1); break;
        default: ew_range<IONP_UINT64, 0>(A, e0, e1); break;  // not planned as flat (other loop dtypes)
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
        if (C.direct) ((T *)A.out)[ob] = fadd(fseed<T>(A), res);
        else ((T *)(partials + C.part_off))[g] = res;
    }
}

template <typename T> __device__ void fadd_fold(const BatchCall &C, const char *partials) {
    const RedArgs &A = C.red;
    const int64_t per_out = A.rsize * A.nwin;
    for (int64_t o = threadIdx.x; o < A.nout; o += blockDim.x) {
        const T *p = (const T *)(partials + C.part_off) + o * per_out;
        T acc = fseed<T>(A);
        for (int64_t i = 0; i < per_out; i++) acc = fadd(acc, p[i]);
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((T *)A.out)[ob] = acc;
    }
}

template <typename T> __device__ void fadd_narrow(const BatchCall &C, int64_t t) {
    const RedArgs &A = C.red;
    const int64_t o = t * kNarrowOutputs + threadIdx.x;
    if (o >= A.nout) return;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    ((T *)A.out)[ob] = narrow_sum<T, 8>(A, ib);  // few registers: the persistent kernel runs every tile kind
}

// lds: PwScratch::racc, 256 elements per wavefront
template <typename T> __device__ void fadd_narrow_wave(const BatchCall &C, int64_t t, double *lds) {
    const RedArgs &A = C.red;
    const int64_t o = t * (kBatchThreads / 64) + (threadIdx.x >> 6);
    if (o >= A.nout) return;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    const T acc = narrow_wave_sum<T, 4>(A, ib, (T *)lds + (threadIdx.x >> 6) * 256);  // few registers: every tile kind shares the kernel
    if ((threadIdx.x & 63) == 0) ((T *)A.out)[ob] = acc;
}

template <typename T> __device__ void fadd_small(const BatchCall &C, int64_t t) {
    constexpr int LT = FDT<T>::v;
    const RedArgs &A = C.red;
    const int64_t o = t * kNarrowOutputs + threadIdx.x;
    if (o >= A.nout) return;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    const int L = (int)A.wsize;
    T acc = fseed<T>(A);
    for (int64_t r = 0; r < A.rsize; r++) {
        const int64_t base = ib + r_offset(A, r);
        acc = fadd(acc, pw_leaf<T>([&](int i) { return ld<LT>(A.in_dt, A.in, base + w_offset(A, i)); }, L));
    }
    ((T *)A.out)[ob] = acc;
}

__device__ inline int64_t group_off(const RedArgs &A, int64_t q) {
    if (A.nr == 0) return w_offset(A, q);
    const int64_t r = q / A.wsize, w = q - r * A.wsize;
    return w_offset(A, w) + r_offset(A, r);
}

template <int LT> __device__ void iadd_small(const BatchCall &C, int64_t t) {
    using ST = typename dt_type<LT>::T;
    const RedArgs &A = C.red;
    const int64_t o = t * kNarrowOutputs + threadIdx.x;
    if (o >= A.nout) return;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    // order-free: eight loads in flight, then their adds
    constexpr int U = 8;
    uint64_t acc = 0;
    int64_t q = 0;
    for (; q + U <= A.group; q += U) {
        uint64_t v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = (uint64_t)ld<LT>(A.in_dt, A.in, ib + group_off(A, q + u));
#pragma unroll
        for (int u = 0; u < U; u++) acc += v[u];
    }
    for (; q < A.group; q++) acc += (uint64_t)ld<LT>(A.in_dt, A.in, ib + group_off(A, q));
    ((ST *)A.out)[ob] = int_result<LT>(acc + (A.has_init ? A.init_bits : 0));
}

// one output per wavefront (kBatchThreads / 64 per tile): lanes take
// group elements lane, lane + 64, ... (four loads in flight), then a
// wavefront sum; the add wraps modulo 2^64, so any order is exact
template <int LT> __device__ void iadd_wave(const BatchCall &C, int64_t t) {
    using ST = typename dt_type<LT>::T;
    const RedArgs &A = C.red;
    const int64_t o = t * (kBatchThreads / 64) + (threadIdx.x >> 6);
    if (o >= A.nout) return;
    const int lane = threadIdx.x & 63;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    constexpr int U = 4;
    uint64_t acc = 0;
    int64_t q = lane;
    for (; q + (U - 1) * 64 < A.group; q += U * 64) {
        uint64_t v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = (uint64_t)ld<LT>(A.in_dt, A.in, ib + group_off(A, q + u * 64));
#pragma unroll
        for (int u = 0; u < U; u++) acc += v[u];
    }
    for (; q < A.group; q += 64) acc += (uint64_t)ld<LT>(A.in_dt, A.in, ib + group_off(A, q));
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
    if (lane == 0) ((ST *)A.out)[ob] = int_result<LT>(acc + (A.has_init ? A.init_bits : 0));
}

template <int LT> __device__ void iadd_tile(const BatchCall &C, char *partials, int64_t g, uint64_t *red) {
    const RedArgs &A = C.red;
    const int64_t o = g / C.nseg, sg = g - o * C.nseg;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    const int64_t q0 = sg * C.tile, q1 = min(A.group, q0 + C.tile);
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
        if (C.direct) ((ST *)A.out)[ob] = int_result<LT>(red[0] + (A.has_init ? A.init_bits : 0));
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
        ((ST *)A.out)[ob] = int_result<LT>(acc + (A.has_init ? A.init_bits : 0));
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

// f(std::integral_constant<int, LT>) for an integer / bool output dtype
template <typename F> __device__ inline void int_dispatch(int dt, const F &f) {
    switch (dt) {
        case IONP_BOOL: f(std::integral_constant<int, IONP_BOOL>()); break;
        case IONP_INT8: f(std::integral_constant<int, IONP_INT8>()); break;
        case IONP_UINT8: f(std::integral_constant<int, IONP_UINT8>()); break;
        case IONP_INT16: f(std::integral_constant<int, IONP_INT16>()); break;
        case IONP_UINT16: f(std::integral_constant<int, IONP_UINT16>()); break;
        case IONP_INT32: f(std::integral_constant<int, IONP_INT32>()); break;
        case IONP_UINT32: f(std::integral_constant<int, IONP_UINT32>()); break;
        case IONP_INT64: f(std::integral_constant<int, IONP_INT64>()); break;
        default: f(std::integral_constant<int, IONP_UINT64>()); break;
    }
}

// Tiles [0, nserial) -- those of the calls with long serial chains, which
// come first -- are handed out one at a time (counters[0]); the rest
// `claim` consecutive tiles per atomic (counters[1]), so a workgroup
// never runs two long chains back to back.
__global__ __launch_bounds__(kBatchThreads) void batch_persistent(
    const BatchCall *__restrict__ calls, const int32_t *__restrict__ tile_map, uint32_t ntiles, uint32_t *counters,
    char *partials, uint32_t nserial, uint32_t claim) {
    __shared__ PwScratch S;
    __shared__ uint64_t ired[kBatchThreads];
    __shared__ int32_t sh_call, sh_local;
    __shared__ BatchCall sC;  // the current call's record: field reads are LDS reads
    static_assert(sizeof(BatchCall) % 8 == 0, "call record copied in 8-byte words");
    int32_t cur = -1;
    uint32_t t_next = 0, t_end = 0;  // thread 0: the claimed range not yet run
    bool past_serial = false;
    for (;;) {
        if (threadIdx.x == 0) {
            if (t_next >= t_end) {
                if (!past_serial) {
                    t_next = atomicAdd(&counters[0], 1u);
                    t_end = t_next + 1;
                    past_serial = t_next >= nserial;
                }
                if (past_serial) {
                    t_next = nserial + atomicAdd(&counters[1], claim);
                    t_end = t_next + claim;
                }
            }
            const uint32_t t = t_next++;
            sh_call = t < ntiles ? tile_map[2 * (size_t)t] : -1;
            sh_local = t < ntiles ? tile_map[2 * (size_t)t + 1] : 0;
        }
        __syncthreads();
        // uniform across the workgroup: the call record is read through
        // scalar loads
        const int32_t c = __builtin_amdgcn_readfirstlane(sh_call);
        if (c < 0) break;  // every tile has been handed out
        const int64_t lt = (int64_t)__builtin_amdgcn_readfirstlane(sh_local);
        if (c != cur) {  // consecutive tiles of one call reuse the copy
            const uint64_t *src = (const uint64_t *)&calls[c];
            for (int i = threadIdx.x; i < (int)(sizeof(BatchCall) / 8); i += blockDim.x) ((uint64_t *)&sC)[i] = src[i];
            __syncthreads();
            cur = c;
        }
        const BatchCall &C = sC;
        switch (C.kind) {
            case BK_ELEMENTWISE: ew_tile(C.ew, lt, C.tile); break;
            case BK_EW_FLAT: ew_flat_tile(C.ew, lt, C.tile); break;
            case BK_EW_SIZED: ew_sized_tile(C.ew, lt, C.tile); break;
            case BK_FADD_WIDE:
                if (C.red.out_dt == IONP_FLOAT64) fadd_window<double>(C, partials, lt, S);
                else fadd_window<float>(C, partials, lt, S);
                break;
            case BK_FADD_SMALL:
                if (C.red.out_dt == IONP_FLOAT64) fadd_small<double>(C, lt);
                else fadd_small<float>(C, lt);
                break;
            case BK_FADD_NARROW:
                if (C.red.out_dt == IONP_FLOAT64) fadd_narrow<double>(C, lt);
                else fadd_narrow<float>(C, lt);
                break;
            case BK_FADD_NARROW_WAVE:
                if (C.red.out_dt == IONP_FLOAT64) fadd_narrow_wave<double>(C, lt, S.racc);
                else fadd_narrow_wave<float>(C, lt, S.racc);
                break;
            case BK_IADD:
                int_dispatch(C.red.out_dt, [&](auto lt_) { iadd_tile<decltype(lt_)::value>(C, partials, lt, ired); });
                break;
            case BK_IADD_SMALL:
                int_dispatch(C.red.out_dt, [&](auto lt_) { iadd_small<decltype(lt_)::value>(C, lt); });
                break;
            case BK_IADD_WAVE:
                int_dispatch(C.red.out_dt, [&](auto lt_) { iadd_wave<decltype(lt_)::value>(C, lt); });
                break;
            default: fill_zero_tile(C, lt); break;
        }
        __syncthreads();
    }
}

// Second launch: calls whose outputs span several tiles fold their partials
// (float: acc = 0; acc = acc + partial in window order; integer: any order).
// Stream order makes every partial of the first launch visible.
__global__ __launch_bounds__(kBatchThreads) void batch_fold(const BatchCall *__restrict__ calls,
                                                            const int32_t *__restrict__ fold_list, int32_t nfold,
                                                            const char *partials) {
    for (int32_t f = blockIdx.x; f < nfold; f += gridDim.x) {
        const BatchCall &C = calls[fold_list[f]];
        if (C.kind == BK_FADD_WIDE) {
            if (C.red.out_dt == IONP_FLOAT64) fadd_fold<double>(C, partials);
            else fadd_fold<float>(C, partials);
        } else {
            int_dispatch(C.red.out_dt, [&](auto lt_) { iadd_fold<decltype(lt_)::value>(C, partials); });
        }
    }
}

ionp_status plan_call_tiles(const ionp_call &c, const ionp_batch_config &cfg, BatchCall *B, int64_t *bytes);

// innermost coalesced dim of stride 1 or 0 in every input and 1 in the
// output, rank <= 4, a loop dtype the vector tiles instantiate
// one coalesced dim, every operand contiguous and aligned to 4 of its elements
static bool flat_ok(const EwArgs &E) {
    if (E.ndim != 1 || (uintptr_t)E.out % (uintptr_t)(kVecV * itemsize(E.out_dt))) return false;
    for (int k = 0; k < E.nin; k++)
        if (E.in_st[k][0] != 1 || (uintptr_t)E.in[k] % (uintptr_t)(kVecV * itemsize(E.in_dt[k]))) return false;
    return true;
}

static bool vec_ok(const EwArgs &E) {
    if (E.ndim < 1 || E.ndim > 4 || E.mixed_su || E.out_st[0] != 1) return false;
    if (E.loop_dt != IONP_FLOAT64 && E.loop_dt != IONP_FLOAT32 && E.loop_dt != IONP_INT64) return false;
    for (int k = 0; k < E.nin; k++)
        if (E.in_st[k][0] != 1 && E.in_st[k][0] != 0) return false;
    return true;
}

// The sized tile path applies: a binary call or where, loop dtype float64 /
// float32 / int64, inputs of 4- or 8-byte items (where's condition bool), at
// most 4 dims, fewer than 2^31 elements and every operand's element offsets
// in (-2^31, 2^31).
static bool sized_ok(const EwArgs &E) {
    if (E.loop_dt != IONP_FLOAT64 && E.loop_dt != IONP_FLOAT32 && E.loop_dt != IONP_INT64) return false;
    if (E.nin < 2 || E.ndim < 1 || E.ndim > 4 || E.mixed_su || E.total >= (int64_t(1) << 31)) return false;
    for (int k = 0; k < E.nin; k++) {
        const int sz = itemsize(E.in_dt[k]);
        if (E.op == OP_WHERE && k == 0 ? E.in_dt[0] != IONP_BOOL : (sz != 4 && sz != 8)) return false;
    }
    for (int k = 0; k <= E.nin; k++) {
        const int64_t *st = k < E.nin ? E.in_st[k] : E.out_st;
        int64_t span = 0;
        for (int d = 0; d < E.ndim; d++) span += (E.shape[d] - 1) * (st[d] < 0 ? -st[d] : st[d]);
        if (span >= (int64_t(1) << 31)) return false;
    }
    return true;
}

// Trailing size-1 dims (stride 0) up to nd: the same iteration space.
static void pad_dims(EwArgs &E, int nd) {
    for (int d = E.ndim; d < nd; d++) {
        E.shape[d] = 1;
        E.div[d] = FastDiv(1);
        for (int k = 0; k < 3; k++) E.in_st[k][d] = 0;
        E.out_st[d] = 0;
    }
    E.ndim = nd;
}

ionp_status plan_call(const ionp_call &c, const ionp_batch_config &cfg, BatchCall *B, int64_t *bytes) {
    std::memset(B, 0, sizeof *B);
    const int nops = c.kind == IONP_CALL_WHERE ? 3 : (c.kind == IONP_CALL_BINARY ? 2 : 1);
    bool half = c.out.dtype == IONP_FLOAT16;
    for (int k = 0; k < nops; k++) half = half || c.in[k].dtype == IONP_FLOAT16;
    ionp_status st = plan_call_tiles(c, cfg, B, bytes);
    if (st == IONP_OK && half) {  // validated and planned; the tile code has no float16 loops
        B->kind = BK_SIDE;
        B->tiles = 0;
        B->direct = 0;
    }
    return st;
}

ionp_status plan_call_tiles(const ionp_call &c, const ionp_batch_config &cfg, BatchCall *B, int64_t *bytes) {
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
        if (c.kind == IONP_CALL_BINARY && c.op >= IONP_LESS && c.out.dtype != IONP_BOOL)
            return set_error(IONP_ERR_TYPE, "comparison output must be bool");
        bool same = c.out.ndim == t.ndim;
        for (int i = 0; same && i < t.ndim; i++) same = c.out.shape[i] == t.shape[i];
        if (!same) return set_error(IONP_ERR_VALUE, "batch call output has the wrong shape");
        if ((s = plan_elementwise(op, nin, ins, &c.out, loop, &B->ew))) return s;
        B->ew.mixed_su = c.kind == IONP_CALL_BINARY && op >= IONP_LESS &&
                         ((c.in[0].dtype == IONP_INT64 && c.in[1].dtype == IONP_UINT64) ||
                          (c.in[0].dtype == IONP_UINT64 && c.in[1].dtype == IONP_INT64));
        if (B->ew.total >= (1ll << 31)) return set_error(IONP_ERR_UNSUPPORTED, "batch call too large");
        for (int d = 0; d < B->ew.ndim; d++) B->ew.div[d] = FastDiv((uint32_t)B->ew.shape[d]);
        // other calls (broadcast operands) keep the generic tiles: 4-element
        // vectors along their rows measured slower on the uneven batch
        // (1.9-2.2 against 1.8 ms for that subset)
        B->kind = vec_ok(B->ew) && flat_ok(B->ew) ? BK_EW_FLAT : (sized_ok(B->ew) ? BK_EW_SIZED : BK_ELEMENTWISE);
        if (B->kind == BK_EW_SIZED) pad_dims(B->ew, B->ew.ndim <= 2 ? 2 : 4);
        B->tile = cfg.ew_tile;
        B->tiles = (B->ew.total + B->tile - 1) / B->tile;
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
        if (p.nw == 0 && narrow_wave_ok(A)) {
            B->kind = BK_FADD_NARROW_WAVE;
            B->tiles = (p.nout + kBatchThreads / 64 - 1) / (kBatchThreads / 64);
        } else if (p.nw == 0) {
            B->kind = BK_FADD_NARROW;
            B->tiles = (p.nout + kNarrowOutputs - 1) / kNarrowOutputs;
        } else if (pw_small_ok(A)) {
            B->kind = BK_FADD_SMALL;
            B->tiles = (p.nout + kNarrowOutputs - 1) / kNarrowOutputs;
        } else {
            B->kind = BK_FADD_WIDE;
            B->tiles = p.nout * p.rsize * A.nwin;
            B->direct = p.rsize == 1 && A.nwin == 1;
        }
    } else if (int_cols_ok(A)) {
        B->kind = BK_IADD_SMALL;
        B->tiles = (p.nout + kNarrowOutputs - 1) / kNarrowOutputs;
    } else if (p.group <= kIntWaveMax) {
        // a workgroup per output would idle most lanes on short groups
        B->kind = BK_IADD_WAVE;
        B->tiles = (p.nout + kBatchThreads / 64 - 1) / (kBatchThreads / 64);
    } else {
        B->kind = BK_IADD;
        B->tile = cfg.int_seg;
        B->nseg = (p.group + B->tile - 1) / B->tile;
        B->tiles = p.nout * B->nseg;
        B->direct = B->nseg == 1;
    }
    return IONP_OK;
}

// one call through its entry point; no configuration means the built-in
// heuristic (a zero ionp_config), never the learned one: per-call and
// graph batches and the autotuner's timings run fixed kernels
static ionp_status run_one(const ionp_call &c, hipStream_t s, const ionp_config *cfg = nullptr) {
    static const ionp_config heuristic = {0, 0, 0, 0};
    if (!cfg) cfg = &heuristic;
    switch (c.kind) {
        case IONP_CALL_BINARY: return ionp_binary(c.op, &c.in[0], &c.in[1], &c.out, cfg, s);
        case IONP_CALL_WHERE: return ionp_where(&c.in[0], &c.in[1], &c.in[2], &c.out, cfg, s);
        case IONP_CALL_CAST: return ionp_cast(&c.in[0], &c.out, cfg, s);
        default: return ionp_reduce(c.op, &c.in[0], c.axes, c.naxes, c.keepdims, nullptr, &c.out, cfg, s);
    }
}

// GRAPH mode, built on first use: one eager pass of the single-call launch
// sequence records the largest scratch request of each kind; the batch
// allocates its own regions of those sizes; the same sequence is then
// captured with those regions installed, so every captured kernel points
// into memory the batch owns and nothing allocates during the capture.
static ionp_status graph_build(ionp_batch *bt, hipStream_t s) {
    ScratchUse use;
    {
        ScratchRecord rec(&use);
        for (const ionp_call &c : bt->calls) {
            ionp_status st0 = run_one(c, s);
            if (st0) return st0;
        }
    }
    auto up = [](size_t b) { return (std::max<size_t>(b, 256) + 255) / 256 * 256; };
    const size_t bs = up(use.scratch), bg = up(use.stage), bl = up(use.slots);
    if (hipStreamSynchronize(s) != hipSuccess || hipMalloc(&bt->d_graph_scratch, bs + bg + bl) != hipSuccess) {
        bt->d_graph_scratch = nullptr;
        return set_error(IONP_ERR_DEVICE, "out of device memory for the graph's scratch");
    }
    ScratchArena &A = bt->graph_arena;
    A = ScratchArena();
    A.scratch = bt->d_graph_scratch;
    A.stage = bt->d_graph_scratch + bs;
    A.slots = bt->d_graph_scratch + bs + bg;
    A.size.scratch = bs;
    A.size.stage = bg;
    A.size.slots = bl;
    (void)num_cus();
    // capture on a private stream (the legacy default stream can not be
    // captured); the graph is then launched on the caller's stream
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return set_error(IONP_ERR_DEVICE, "hipStreamCreate failed");
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return set_error(IONP_ERR_DEVICE, "hipStreamBeginCapture failed");
    }
    ionp_status st = IONP_OK;
    {
        ScratchScope scope(&A);
        for (const ionp_call &c : bt->calls)
            if (!st) st = run_one(c, cs);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (!st && A.short_) st = set_error(IONP_ERR_DEVICE, "graph capture asked for more scratch than the eager pass");
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
    return IONP_OK;
}

}  // namespace ionp

using namespace ionp;

extern "C" {

ionp_status ionp_batch_create(const ionp_call *calls, int32_t n, void *stream, ionp_batch **out) {
    return ionp_batch_create_ex(calls, n, nullptr, stream, out);
}

ionp_status ionp_batch_create_ex(const ionp_call *calls, int32_t n, const ionp_batch_config *config, void *stream,
                                 ionp_batch **out) {
    if (!out || (n > 0 && !calls) || n < 0) return set_error(IONP_ERR_VALUE, "bad batch arguments");
    *out = nullptr;
    ionp_batch_config cfg = {kEwTileDefault, kIntSegDefault, kClaimDefault, 0};
    if (config) {
        if (config->ew_tile) cfg.ew_tile = config->ew_tile;
        if (config->int_seg) cfg.int_seg = config->int_seg;
        if (config->claim) cfg.claim = config->claim;
    }
    if (cfg.claim < 1 || cfg.claim > 64) return set_error(IONP_ERR_VALUE, "ionp_batch_config: claim must be in [1, 64]");
    // elementwise tiles: whole 4-element vectors of whole 256-thread steps
    if (cfg.ew_tile < 1024 || cfg.ew_tile > (1 << 24) || cfg.ew_tile % 1024 || cfg.int_seg < 256 ||
        cfg.int_seg > (1 << 24))
        return set_error(IONP_ERR_VALUE, "ionp_batch_config: ew_tile must be a multiple of 1024 in [1024, 2^24], "
                                         "int_seg in [256, 2^24]");
    std::vector<BatchCall> planned(n > 0 ? n : 1), table(n > 0 ? n : 1);
    std::vector<int64_t> begin(n + 1, 0), serial(n > 0 ? n : 1, 0);
    std::vector<int32_t> fold_list, order(n), side;
    int64_t nserial = 0;
    int64_t bytes = 0, part = 0;
    size_t scratch_need = 0;
    for (int32_t i = 0; i < n; i++) {
        int64_t b = 0;
        ionp_status s = plan_call(calls[i], cfg, &planned[i], &b);
        if (s) {
            return set_error(s, "batch call " + std::to_string(i) + ": " + last_error());
        }
        bytes += b;
        const BatchCall &B = planned[i];
        // length of the dependent chain one thread runs in a tile of this call
        serial[i] = B.kind == BK_FADD_NARROW || B.kind == BK_FADD_NARROW_WAVE ? B.red.rsize
                    : B.kind == BK_FADD_SMALL ? B.red.rsize * B.red.wsize
                    : B.kind == BK_IADD_SMALL ? B.red.group
                    : B.kind == BK_IADD_WAVE  ? B.red.group / 64
                                              : 0;
        order[i] = i;
        if (B.kind == BK_SIDE) side.push_back(i);
    }
    // tiles are handed out in table order: long serial tiles first, so they do
    // not finish the launch alone (the order of calls never changes a result)
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return serial[a] > serial[b]; });
    for (int32_t i = 0; i < n; i++) {
        table[i] = planned[order[i]];
        BatchCall &B = table[i];
        if ((B.kind == BK_FADD_WIDE || B.kind == BK_IADD) && !B.direct) {
            B.part_off = part;
            part += ((B.tiles * 8) + 255) / 256 * 256;
            scratch_need = std::max(scratch_need, (size_t)(B.tiles * 16));
            fold_list.push_back(i);
        }
        begin[i + 1] = begin[i] + B.tiles;
        if (serial[order[i]] > kSerialChain) nserial = begin[i + 1];
    }
    ionp_batch *bt = new ionp_batch();
    bt->n = n;
    bt->ntiles = begin[n];
    bt->bytes = bytes;
    bt->part_bytes = part;
    bt->calls.assign(calls, calls + n);
    bt->side = side;
    bt->claim = cfg.claim;
    bt->nserial = nserial;
    bt->graph_ready = false;
    bt->scratch_need = scratch_need;
    bt->d_graph_scratch = nullptr;
    hipStream_t s = (hipStream_t)stream;
    bt->nfold = (int32_t)fold_list.size();
    if (begin[n] >= (1ll << 31)) {
        delete bt;
        return set_error(IONP_ERR_UNSUPPORTED, "batch has 2^31 tiles or more");
    }
    std::vector<int32_t> tmap(2 * (size_t)std::max<int64_t>(begin[n], 1));
    for (int32_t i = 0; i < n; i++)
        for (int64_t t = begin[i]; t < begin[i + 1]; t++) {
            tmap[2 * t] = i;
            tmap[2 * t + 1] = (int32_t)(t - begin[i]);
        }
    bool ok = hipMalloc(&bt->d_calls, sizeof(BatchCall) * table.size()) == hipSuccess &&
              hipMalloc(&bt->d_tile_map, sizeof(int32_t) * tmap.size()) == hipSuccess &&
              hipMalloc(&bt->d_counters, sizeof(uint32_t) * 64) == hipSuccess &&
              hipMalloc(&bt->d_fold_list, sizeof(int32_t) * (fold_list.size() + 1)) == hipSuccess &&
              hipMalloc(&bt->d_partials, part > 0 ? part : 256) == hipSuccess;
    ok = ok && hipMemcpyAsync(bt->d_calls, table.data(), sizeof(BatchCall) * table.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
         hipMemcpyAsync(bt->d_tile_map, tmap.data(), sizeof(int32_t) * tmap.size(), hipMemcpyHostToDevice, s) == hipSuccess &&
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
    FpeOff plain;  // batches run the plain kernels whatever the thread's IEEE flag tracking
    if (mode == IONP_BATCH_PER_CALL) {
        for (const ionp_call &c : bt->calls) {
            ionp_status st = run_one(c, s);
            if (st) return st;
        }
        return IONP_OK;
    }
    if (mode == IONP_BATCH_GRAPH) {
        if (!bt->graph_ready) {
            ionp_status st0 = graph_build(bt, s);
            if (st0) return st0;
        }
        return hipGraphLaunch(bt->graph, s) == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "hipGraphLaunch failed");
    }
    if (mode != IONP_BATCH_PERSISTENT) return set_error(IONP_ERR_VALUE, "unknown batch mode");
    for (int32_t i : bt->side) {  // float16 calls: their own launches, in stream order with the kernel below
        ionp_status st = run_one(bt->calls[i], s);
        if (st) return st;
    }
    if (bt->n == 0 || bt->ntiles == 0) return IONP_OK;
    hipLaunchKernelGGL(batch_zero, dim3(1), dim3(64), 0, s, bt->d_counters, (int64_t)64);
    // persistent grid: as many workgroups as fit on the chip at once (the
    // kernel's register and LDS use decide), never more than the tiles
    static int per_cu = 0;
    if (!per_cu) {
        int b = 0;
        per_cu = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batch_persistent, kBatchThreads, 0) == hipSuccess && b > 0
                     ? b
                     : 2;
    }
    int64_t grid = (int64_t)num_cus() * per_cu;
    if (grid > bt->ntiles) grid = bt->ntiles;
    hipLaunchKernelGGL(batch_persistent, dim3((unsigned)grid), dim3(kBatchThreads), 0, s, bt->d_calls, bt->d_tile_map,
                       (uint32_t)bt->ntiles, bt->d_counters, bt->d_partials, (uint32_t)bt->nserial, (uint32_t)bt->claim);
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
    FpeOff plain;  // timings of the plain kernels
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

ionp_status ionp_batch_autotune(const ionp_call *calls, int32_t n, const ionp_batch_config *cands, int32_t ncand,
                                int32_t reps, float *times_us, void *stream) {
    if (n < 0 || ncand <= 0 || reps <= 0 || !cands || !times_us || (n > 0 && !calls))
        return set_error(IONP_ERR_VALUE, "bad batch autotune arguments");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return set_error(IONP_ERR_DEVICE, "hipEventCreate failed");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(IONP_ERR_DEVICE, "hipEventCreate failed");
    }
    ionp_status st = IONP_OK;
    for (int32_t j = 0; j < ncand && !st; j++) {
        times_us[j] = -1.0f;
        ionp_batch *bt = nullptr;
        if (ionp_batch_create_ex(calls, n, &cands[j], stream, &bt) != IONP_OK) continue;  // unusable schedule
        st = ionp_batch_run(bt, IONP_BATCH_PERSISTENT, stream);  // warm
        for (int32_t r = 0; r < reps && !st; r++) {
            (void)hipEventRecord(e0, s);
            st = ionp_batch_run(bt, IONP_BATCH_PERSISTENT, stream);
            (void)hipEventRecord(e1, s);
            if (!st && hipEventSynchronize(e1) != hipSuccess) st = set_error(IONP_ERR_DEVICE, "batch autotune run failed");
            float ms = 0.0f;
            if (!st && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && (times_us[j] < 0.0f || ms * 1000.0f < times_us[j]))
                times_us[j] = ms * 1000.0f;
        }
        ionp_batch_destroy(bt);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

void ionp_batch_destroy(ionp_batch *bt) {
    if (!bt) return;
    if (bt->graph_ready) (void)hipGraphExecDestroy(bt->graph);
    if (bt->d_calls) (void)hipFree(bt->d_calls);
    if (bt->d_tile_map) (void)hipFree(bt->d_tile_map);
    if (bt->d_counters) (void)hipFree(bt->d_counters);
    if (bt->d_fold_list) (void)hipFree(bt->d_fold_list);
    if (bt->d_partials) (void)hipFree(bt->d_partials);
    if (bt->d_graph_scratch) (void)hipFree(bt->d_graph_scratch);
    delete bt;
}

}  // extern "C"
