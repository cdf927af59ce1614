// rows_select_cast.hip -- the rows layout of ew_rows for where_select and
// casts.
//
// Replaces where_select's select loop (ionp-core/src/sort.rs:1885-1913) and
// Buffer::cast_to's per-element conversion (ionp-core/src/buffer.rs:81-210)
// for "rows x inner" layouts: the innermost coalesced axis is contiguous in
// the output and contiguous, reversed (stride -1) or broadcast (stride 0)
// in every input.
//   ew_rows_where<T> - a bool condition and two value operands of the
//                      output's dtype: a select of bit patterns, so one
//                      kernel per item size serves every dtype.
//   ew_rows_cast<LI, LO> - any pair of non-float16 dtypes, each element
//                      converted by cvt exactly as the generic kernel does
//                      (a bool input is read as 0 / 1 first).
// Each thread moves 2 chunks of V elements (V * itemsize = 16 bytes for the
// widest operand), interleaved so that each wavefront instruction covers a
// contiguous span, all loaded before any element is selected or converted;
// one tile of 4 * V * blockDim elements per workgroup.  Everything else (IEEE flags
// tracked, float16, a non-bool condition, mixed value dtypes, misaligned
// views, other strides) stays on the generic kernel.
#include "elementwise.h"

#include "device_ops.h"

namespace ionp {

template <typename T, int N> using VecT = T __attribute__((ext_vector_type(N)));

struct RowArgs3 {
    int64_t inner, rows, tiles_per_row;
    FastDiv tpr;
    int32_t nouter;
    int64_t oshape[kMaxKDims];
    FastDiv odiv[kMaxKDims];
    int64_t ost[3][kMaxKDims], ost_out[kMaxKDims];
    int64_t inner_st[3];  // 0 (broadcast), 1 or -1 (a reversed axis)
    const char *in[3];
    char *out;
    int32_t nt_loads;     // stream the inputs with nontemporal loads
};

// element offsets of row `row` in the operands and the output
__device__ inline void row_offsets(const RowArgs3 &R, uint32_t row, int nin, int64_t *r, int64_t &ro) {
    for (int k = 0; k < 3; k++) r[k] = 0;
    ro = 0;
    uint32_t rem = row;
    for (int d = 0; d < R.nouter; d++) {
        uint32_t c;
        if (d + 1 < R.nouter) {
            const uint32_t q = R.odiv[d].div(rem);
            c = rem - q * (uint32_t)R.oshape[d];
            rem = q;
        } else {
            c = rem;
        }
        for (int k = 0; k < nin; k++) r[k] += (int64_t)c * R.ost[k][d];
        ro += (int64_t)c * R.ost_out[d];
    }
}

// Each thread moves U chunks of V elements, chunk j at elements
// (j * blockDim + thread) * V of its tile, V = 16 / the widest item size:
// every load or store instruction of a wavefront covers one contiguous
// span (64 x 16 bytes for the widest operand, 64 x V x itemsize for the
// others), and all U chunks of every operand are loaded before any element
// is selected or converted.
// Measured on MI355X, 8192 x 8192: where(cond, A, c) float64 U 1 / 2 / 4
// 209 / 190 / 225 us, float32 109 / 104 / 118 us; A.astype(float32) 147 /
// 133 / 137 us, float32 -> float64 117 / 114 / 116 us.
constexpr int kRowsU = 2;

template <typename T, int V> __device__ inline VecT<T, V> vload(const T *p, bool nt) {
    return nt ? __builtin_nontemporal_load((const VecT<T, V> *)p) : *(const VecT<T, V> *)p;
}

// elements i .. i + V - 1 of a row read at stride st (1, or -1: a reversed
// axis, whose V elements are the aligned vector ending at p - i, reversed)
template <typename T, int V> __device__ inline VecT<T, V> vload_dir(const T *p, int64_t i, int64_t st, bool nt) {
    if (st > 0) return vload<T, V>(p + i, nt);
    const VecT<T, V> v = vload<T, V>(p - i - (V - 1), nt);
    VecT<T, V> r;
#pragma unroll
    for (int k = 0; k < V; k++) r[k] = v[V - 1 - k];
    return r;
}

// chunks per thread of the where kernel: fewer for narrow items, whose
// select unpacks every item into a register
template <typename T> constexpr int where_u() { return sizeof(T) >= 2 ? kRowsU : 1; }

template <typename T> __global__ __launch_bounds__(256) void ew_rows_where(const RowArgs3 R) {
    constexpr int V = 16 / (int)sizeof(T), U = where_u<T>();
    const int64_t tile = (int64_t)blockDim.x * V * U;
    const uint32_t ntiles = (uint32_t)(R.rows * R.tiles_per_row);
    const bool nt = R.nt_loads;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t row = R.tpr.div(t);
        const int64_t i0 = (int64_t)(t - row * (uint32_t)R.tiles_per_row) * tile;
        int64_t r[3], ro;
        row_offsets(R, row, 3, r, ro);
        const uint8_t *pc = (const uint8_t *)R.in[0] + r[0];
        const T *px = (const T *)R.in[1] + r[1];
        const T *py = (const T *)R.in[2] + r[2];
        T *po = (T *)R.out + ro;
        const uint8_t c0 = R.inner_st[0] ? 0 : pc[0];
        const T x0 = R.inner_st[1] ? T(0) : px[0], y0 = R.inner_st[2] ? T(0) : py[0];
        VecT<uint8_t, V> c[U];
        VecT<T, V> x[U], y[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int64_t i = i0 + ((int64_t)j * blockDim.x + threadIdx.x) * V;
            if (i + V <= R.inner) {
                if (R.inner_st[0]) c[j] = vload_dir<uint8_t, V>(pc, i, R.inner_st[0], nt);
                if (R.inner_st[1]) x[j] = vload_dir<T, V>(px, i, R.inner_st[1], nt);
                if (R.inner_st[2]) y[j] = vload_dir<T, V>(py, i, R.inner_st[2], nt);
            }
        }
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int64_t i = i0 + ((int64_t)j * blockDim.x + threadIdx.x) * V;
            if (i + V <= R.inner) {
                VecT<T, V> o;
#pragma unroll
                for (int k = 0; k < V; k++) {
                    const bool ck = (R.inner_st[0] ? c[j][k] : c0) != 0;
                    o[k] = ck ? (R.inner_st[1] ? x[j][k] : x0) : (R.inner_st[2] ? y[j][k] : y0);
                }
                __builtin_nontemporal_store(o, (VecT<T, V> *)(po + i));
            } else {
                for (int64_t e = i; e < R.inner && e < i + V; e++)
                    po[e] = pc[e * R.inner_st[0]] != 0 ? px[e * R.inner_st[1]] : py[e * R.inner_st[2]];
            }
        }
    }
}

template <int LI, int LO> __global__ __launch_bounds__(256) void ew_rows_cast(const RowArgs3 R) {
    using TI = typename dt_type<LI>::T;
    using TO = typename dt_type<LO>::T;
    constexpr int V = 16 / (sizeof(TI) > sizeof(TO) ? (int)sizeof(TI) : (int)sizeof(TO)), U = kRowsU;
    const int64_t tile = (int64_t)blockDim.x * V * U;
    const uint32_t ntiles = (uint32_t)(R.rows * R.tiles_per_row);
    const bool nt = R.nt_loads;
    auto conv = [](TI v) -> TO {
        if constexpr (LI == IONP_BOOL) return cvt<LO>((uint8_t)(v != 0));
        else return cvt<LO>(v);
    };
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t row = R.tpr.div(t);
        const int64_t i0 = (int64_t)(t - row * (uint32_t)R.tiles_per_row) * tile;
        int64_t r[3], ro;
        row_offsets(R, row, 1, r, ro);
        const TI *pa = (const TI *)R.in[0] + r[0];
        TO *po = (TO *)R.out + ro;
        VecT<TI, V> a[U];
        if (R.inner_st[0]) {
#pragma unroll
            for (int j = 0; j < U; j++) {
                const int64_t i = i0 + ((int64_t)j * blockDim.x + threadIdx.x) * V;
                if (i + V <= R.inner) a[j] = vload_dir<TI, V>(pa, i, R.inner_st[0], nt);
            }
        }
        const TO b = R.inner_st[0] ? TO(0) : conv(pa[0]);
#pragma unroll
        for (int j = 0; j < U; j++) {
            const int64_t i = i0 + ((int64_t)j * blockDim.x + threadIdx.x) * V;
            if (i + V <= R.inner) {
                VecT<TO, V> o;
#pragma unroll
                for (int k = 0; k < V; k++) o[k] = R.inner_st[0] ? conv(a[j][k]) : b;
                __builtin_nontemporal_store(o, (VecT<TO, V> *)(po + i));
            } else {
                for (int64_t e = i; e < R.inner && e < i + V; e++) po[e] = conv(pa[e * R.inner_st[0]]);
            }
        }
    }
}

typedef void (*Row3Fn)(const RowArgs3);

static Row3Fn where_kernel(int isz) {
    switch (isz) {
        case 1: return ew_rows_where<uint8_t>;
        case 2: return ew_rows_where<uint16_t>;
        case 4: return ew_rows_where<uint32_t>;
        case 8: return ew_rows_where<uint64_t>;
        default: return nullptr;
    }
}

template <int LI> static Row3Fn cast_kernel_from(int lo) {
    switch (lo) {
        case IONP_BOOL: return ew_rows_cast<LI, IONP_BOOL>;
        case IONP_INT8: return ew_rows_cast<LI, IONP_INT8>;
        case IONP_UINT8: return ew_rows_cast<LI, IONP_UINT8>;
        case IONP_INT16: return ew_rows_cast<LI, IONP_INT16>;
        case IONP_UINT16: return ew_rows_cast<LI, IONP_UINT16>;
        case IONP_INT32: return ew_rows_cast<LI, IONP_INT32>;
        case IONP_UINT32: return ew_rows_cast<LI, IONP_UINT32>;
        case IONP_INT64: return ew_rows_cast<LI, IONP_INT64>;
        case IONP_UINT64: return ew_rows_cast<LI, IONP_UINT64>;
        case IONP_FLOAT32: return ew_rows_cast<LI, IONP_FLOAT32>;
        case IONP_FLOAT64: return ew_rows_cast<LI, IONP_FLOAT64>;
        default: return nullptr;
    }
}

static Row3Fn cast_kernel(int li, int lo) {
    switch (li) {
        case IONP_BOOL: return cast_kernel_from<IONP_BOOL>(lo);
        case IONP_INT8: return cast_kernel_from<IONP_INT8>(lo);
        case IONP_UINT8: return cast_kernel_from<IONP_UINT8>(lo);
        case IONP_INT16: return cast_kernel_from<IONP_INT16>(lo);
        case IONP_UINT16: return cast_kernel_from<IONP_UINT16>(lo);
        case IONP_INT32: return cast_kernel_from<IONP_INT32>(lo);
        case IONP_UINT32: return cast_kernel_from<IONP_UINT32>(lo);
        case IONP_INT64: return cast_kernel_from<IONP_INT64>(lo);
        case IONP_UINT64: return cast_kernel_from<IONP_UINT64>(lo);
        case IONP_FLOAT32: return cast_kernel_from<IONP_FLOAT32>(lo);
        case IONP_FLOAT64: return cast_kernel_from<IONP_FLOAT64>(lo);
        default: return nullptr;
    }
}

bool try_rows_select_cast(const EwArgs &A, const ionp_config *cfg, hipStream_t s, ionp_status *st) {
    if (A.op != OP_WHERE && A.op != OP_CAST) return false;
    if (A.fpe) return false;  // flags tracked: the generic kernel's flag variant
    if (A.ndim > 4 || A.shape[0] < 64 || A.out_st[0] != 1) return false;
    const int nin = A.op == OP_WHERE ? 3 : 1;
    if (A.nin != nin) return false;
    for (int k = 0; k < nin; k++)  // contiguous, broadcast or reversed along the inner axis
        if (A.in_st[k][0] != 0 && A.in_st[k][0] != 1 && A.in_st[k][0] != -1) return false;
    int isz[3] = {0, 0, 0}, V, U = kRowsU;
    const int osz = itemsize(A.out_dt);
    Row3Fn fn;
    if (A.op == OP_WHERE) {
        // bool values are read as 0 / 1 by the generic kernel: not a bit copy
        if (A.in_dt[0] != IONP_BOOL || A.in_dt[1] != A.out_dt || A.in_dt[2] != A.out_dt || A.out_dt == IONP_BOOL)
            return false;
        isz[0] = 1;
        isz[1] = isz[2] = osz;
        V = 16 / osz;
        U = osz >= 2 ? kRowsU : 1;  // where_u
        fn = where_kernel(osz);
    } else {
        isz[0] = itemsize(A.in_dt[0]);
        V = 16 / (isz[0] > osz ? isz[0] : osz);
        fn = cast_kernel(A.in_dt[0], A.out_dt);
    }
    if (!fn) return false;
    // every contiguous row must start on its vector access size (V items);
    // a reversed row must end there (its first element is the last of a vector)
    auto aligned = [&](const char *p, int sz, const int64_t *st, int64_t inner_st) {
        if (inner_st == 0) return true;
        const int64_t vb = (int64_t)V * sz;
        if (((uintptr_t)p + (inner_st < 0 ? (uintptr_t)sz : 0)) % (uintptr_t)vb) return false;
        for (int d = 1; d < A.ndim; d++)
            if ((st[d] * sz) % vb) return false;
        return true;
    };
    for (int k = 0; k < nin; k++)
        if (!aligned(A.in[k], isz[k], A.in_st[k], A.in_st[k][0])) return false;
    if (!aligned(A.out, osz, A.out_st, 1)) return false;
    const int block = (cfg && cfg->block > 0) ? cfg->block : 256;
    RowArgs3 R;
    std::memset(&R, 0, sizeof R);
    R.inner = A.shape[0];
    R.rows = 1;
    R.nouter = A.ndim - 1;
    for (int d = 1; d < A.ndim; d++) {
        R.oshape[d - 1] = A.shape[d];
        R.odiv[d - 1] = FastDiv((uint32_t)A.shape[d]);
        for (int k = 0; k < nin; k++) R.ost[k][d - 1] = A.in_st[k][d];
        R.ost_out[d - 1] = A.out_st[d];
        R.rows *= A.shape[d];
    }
    int64_t in_bytes = 0;
    for (int k = 0; k < nin; k++) {
        R.inner_st[k] = A.in_st[k][0];
        R.in[k] = A.in[k];
        if (A.in_st[k][0]) in_bytes += R.inner * R.rows * isz[k];
    }
    R.out = A.out;
    R.nt_loads = in_bytes > kInfinityCacheBytes;
    const int64_t tile = (int64_t)block * V * U;
    R.tiles_per_row = (R.inner + tile - 1) / tile;
    if (R.rows * R.tiles_per_row >= (1ll << 31) || R.rows >= (1ll << 31)) return false;
    R.tpr = FastDiv((uint32_t)R.tiles_per_row);
    const int64_t ntiles = R.rows * R.tiles_per_row;
    // a workgroup per tile (as ew_rows); a configured blocks_per_cu caps the grid
    int64_t grid = (cfg && cfg->blocks_per_cu > 0) ? (int64_t)cfg->blocks_per_cu * num_cus() : ntiles;
    if (grid > ntiles) grid = ntiles;
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(block), 0, s, R);
    *st = hipGetLastError() == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "kernel launch failed");
    return true;
}

}  // namespace ionp
