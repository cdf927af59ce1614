// elementwise.hip -- gfx950 kernels for binary ufuncs, where and casts.
//
// Replaces the reference's inner loops: binary_elementwise + same_kind
// dispatch (ionp-core/src/ufunc.rs:3539-3600), where_select's select loop
// (ionp-core/src/sort.rs:1885-1913) and Buffer::cast_to
// (ionp-core/src/buffer.rs:210).  The reference walks two NdIter odometers
// per element and materialises casts and broadcasts as copies; here casts
// are fused into the loads, broadcasts are stride-0 reads and the planner
// coalesces the iteration space first.
//
// Three kernels:
//   ew_tiled   - one operand contiguous along an axis other than the output's
//                contiguous one (A + B.T): 64 x 64 tiles staged through LDS.
//   ew_rows    - "rows x inner" layouts whose innermost axis is contiguous
//                in the output and contiguous or broadcast (stride 0) in each
//                input, all operands of one dtype: 16-byte vector loads and
//                stores, several vectors in flight per thread.
//   ew_generic - any strides (rank <= 8 after coalescing), any dtype mix.
// (where and casts over rows layouts: rows_select_cast.hip; any call with
// a transposed input that ew_tiled does not take: tiled_any.hip)
#include "elementwise.h"

#include "device_ops.h"

namespace ionp {

// Generic strided kernel: element e of the coalesced space (dim 0
// innermost) -> per-operand offsets by fast division.  A.total < 2^31.
// FPE: also OR the IEEE flags of the elements into *A.fpe.
template <int LT, bool FPE = false> __global__ __launch_bounds__(256) void ew_generic(const EwArgs A) {
    const int i0 = itemsize(A.in_dt[0]), i1 = itemsize(A.in_dt[1]), i2 = itemsize(A.in_dt[2]);
    const int io = itemsize(A.out_dt);
    const uint32_t total = (uint32_t)A.total;
    uint32_t fl = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
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
        if constexpr (FPE) fl |= ew_element_fpe<LT>(A, A.in[0] + o0 * i0, A.in[1] + o1 * i1, A.in[2] + o2 * i2, A.out + oo * io);
        else ew_element<LT>(A, A.in[0] + o0 * i0, A.in[1] + o1 * i1, A.in[2] + o2 * i2, A.out + oo * io);
    }
    if constexpr (FPE) fpe_commit(A.fpe, fl);
}

// Calls with a float16 operand or output: the generic walk with
// ew_half_elem; FPE: OR the IEEE flags into *A.fpe.
template <bool FPE> __global__ __launch_bounds__(256) void ew_half(const EwArgs A) {
    const int i0 = itemsize(A.in_dt[0]), i1 = itemsize(A.in_dt[1]), i2 = itemsize(A.in_dt[2]);
    const int io = itemsize(A.out_dt);
    const uint32_t total = (uint32_t)A.total;
    uint32_t fl = 0;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
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
        fl |= ew_half_elem(A, A.in[0] + o0 * i0, A.in[1] + o1 * i1, A.in[2] + o2 * i2, A.out + oo * io);
    }
    if constexpr (FPE) fpe_commit(A.fpe, fl);
}

// ---------------------------------------------------------------------------
// rows x inner fast path
// ---------------------------------------------------------------------------
template <typename T, int N> using VecT = T __attribute__((ext_vector_type(N)));

// OPK: compile-time op so each instantiation is a straight-line loop.
template <typename T, int OPK> __device__ inline auto row_apply(T a, T b) {
    if constexpr (OPK >= IONP_LESS) {
        return (uint8_t)cmp_op<T>(OPK, a, b);
    } else {
        constexpr int LT = std::is_same<T, double>::value     ? IONP_FLOAT64
                           : std::is_same<T, float>::value    ? IONP_FLOAT32
                           : std::is_same<T, int64_t>::value  ? IONP_INT64
                           : std::is_same<T, uint64_t>::value ? IONP_UINT64
                           : std::is_same<T, int32_t>::value  ? IONP_INT32
                           : std::is_same<T, uint32_t>::value ? IONP_UINT32
                           : std::is_same<T, int16_t>::value  ? IONP_INT16
                           : std::is_same<T, uint16_t>::value ? IONP_UINT16
                           : std::is_same<T, int8_t>::value   ? IONP_INT8
                                                              : IONP_UINT8;
        return arith_op<LT>(OPK, a, b);
    }
}

// GATHER: inputs whose inner stride is neither 0 nor 1 (e.g. X.T[::2]) are
// read element by element at that stride; the output stays vectorised.
// FPE: also OR the IEEE flags of the elements into *R.fpe.
template <typename T, int OPK, int VEC, int UNROLL, bool GATHER = false, bool FPE = false>
__global__ __launch_bounds__(256) void ew_rows(const RowArgs R) {
    using OT = typename std::conditional<(OPK >= IONP_LESS), uint8_t, T>::type;
    uint32_t fl = 0;
    const T *__restrict__ a = (const T *)R.in[0];
    const T *__restrict__ b = (const T *)R.in[1];
    OT *__restrict__ o = (OT *)R.out;
    const int64_t tile = (int64_t)blockDim.x * VEC * UNROLL;
    const uint32_t ntiles = (uint32_t)(R.rows * R.tiles_per_row);
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t row = R.tpr.div(t);
        const int64_t i0 = (int64_t)(t - row * (uint32_t)R.tiles_per_row) * tile;
        // row base offsets: decompose the row index over the outer dims
        int64_t ra = 0, rb = 0, ro = 0;
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
            ra += (int64_t)c * R.ost[0][d];
            rb += (int64_t)c * R.ost[1][d];
            ro += (int64_t)c * R.ost_out[d];
        }
        const T *pa = a + ra;
        const T *pb = b + rb;
        OT *po = o + ro;
        const int64_t sa = R.inner_st[0], sb = R.inner_st[1];
        const T a0 = sa ? T(0) : pa[0];
        const T b0 = sb ? T(0) : pb[0];
        VecT<T, VEC> va[UNROLL], vb[UNROLL];
        const int64_t sta = R.inner_st[0], stb = R.inner_st[1];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t i = i0 + ((int64_t)u * blockDim.x + threadIdx.x) * VEC;
            if (i + VEC <= R.inner) {
                if constexpr (GATHER) {
                    // stride 2 (x[::2] views): two 16-byte loads, keep the even
                    // elements; other strides element by element
                    if (sa == 2 && R.pair_ok[0]) {
                        const VecT<T, VEC> lo = *(const VecT<T, VEC> *)(pa + 2 * i);
                        const VecT<T, VEC> hi = *(const VecT<T, VEC> *)(pa + 2 * i + VEC);
#pragma unroll
                        for (int k = 0; k < VEC; k++) va[u][k] = k < VEC / 2 ? lo[2 * k] : hi[2 * k - VEC];
                    } else if (sa) {
#pragma unroll
                        for (int k = 0; k < VEC; k++) va[u][k] = pa[(i + k) * sta];
                    }
                    if (stb == 2 && R.pair_ok[1]) {
                        const VecT<T, VEC> lo = *(const VecT<T, VEC> *)(pb + 2 * i);
                        const VecT<T, VEC> hi = *(const VecT<T, VEC> *)(pb + 2 * i + VEC);
#pragma unroll
                        for (int k = 0; k < VEC; k++) vb[u][k] = k < VEC / 2 ? lo[2 * k] : hi[2 * k - VEC];
                    } else if (sb) {
#pragma unroll
                        for (int k = 0; k < VEC; k++) vb[u][k] = pb[(i + k) * stb];
                    }
                } else {
                    // inputs larger than the Infinity Cache stream through
                    // with nontemporal loads (+5% on 1 GiB operands); smaller
                    // ones load normally and can stay cache-resident
                    if (R.nt_loads) {
                        if (sa) va[u] = __builtin_nontemporal_load((const VecT<T, VEC> *)(pa + i));
                        if (sb) vb[u] = __builtin_nontemporal_load((const VecT<T, VEC> *)(pb + i));
                    } else {
                        if (sa) va[u] = *(const VecT<T, VEC> *)(pa + i);
                        if (sb) vb[u] = *(const VecT<T, VEC> *)(pb + i);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const int64_t i = i0 + ((int64_t)u * blockDim.x + threadIdx.x) * VEC;
            if (i + VEC <= R.inner) {
                VecT<OT, VEC> r;
#pragma unroll
                for (int k = 0; k < VEC; k++) {
                    const T x = sa ? va[u][k] : a0, y = sb ? vb[u][k] : b0;
                    r[k] = row_apply<T, OPK>(x, y);
                    if constexpr (FPE) fl |= fpe_arith<T, OPK>(x, y, r[k]);
                }
                __builtin_nontemporal_store(r, (VecT<OT, VEC> *)(po + i));
            } else {
                for (int64_t j = i; j < R.inner && j < i + VEC; j++) {
                    const T x = sa ? pa[j * sta] : a0, y = sb ? pb[j * stb] : b0;
                    po[j] = row_apply<T, OPK>(x, y);
                    if constexpr (FPE) fl |= fpe_arith<T, OPK>(x, y, po[j]);
                }
            }
        }
    }
    if constexpr (FPE) fpe_commit(R.fpe, fl);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int g_num_cus = 0;
int num_cus() {
    if (!g_num_cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&g_num_cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (g_num_cus <= 0) g_num_cus = 256;
    }
    return g_num_cus;
}

static int grid_for(int64_t work_items, int block, const ionp_config *cfg) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    const int64_t need = (work_items + block - 1) / block;
    int64_t g = (int64_t)bpc * num_cus();
    if (g > need) g = need;
    return (int)(g > 0 ? g : 1);
}

typedef void (*GenericFn)(const EwArgs);

static GenericFn generic_kernel(int lt, bool fpe) {
    if (fpe) {
        switch (lt) {
            case IONP_BOOL: return ew_generic<IONP_BOOL, true>;
            case IONP_INT8: return ew_generic<IONP_INT8, true>;
            case IONP_UINT8: return ew_generic<IONP_UINT8, true>;
            case IONP_INT16: return ew_generic<IONP_INT16, true>;
            case IONP_UINT16: return ew_generic<IONP_UINT16, true>;
            case IONP_INT32: return ew_generic<IONP_INT32, true>;
            case IONP_UINT32: return ew_generic<IONP_UINT32, true>;
            case IONP_INT64: return ew_generic<IONP_INT64, true>;
            case IONP_UINT64: return ew_generic<IONP_UINT64, true>;
            case IONP_FLOAT32: return ew_generic<IONP_FLOAT32, true>;
            default: return ew_generic<IONP_FLOAT64, true>;
        }
    }
    switch (lt) {
        case IONP_BOOL: return ew_generic<IONP_BOOL>;
        case IONP_INT8: return ew_generic<IONP_INT8>;
        case IONP_UINT8: return ew_generic<IONP_UINT8>;
        case IONP_INT16: return ew_generic<IONP_INT16>;
        case IONP_UINT16: return ew_generic<IONP_UINT16>;
        case IONP_INT32: return ew_generic<IONP_INT32>;
        case IONP_UINT32: return ew_generic<IONP_UINT32>;
        case IONP_INT64: return ew_generic<IONP_INT64>;
        case IONP_UINT64: return ew_generic<IONP_UINT64>;
        case IONP_FLOAT32: return ew_generic<IONP_FLOAT32>;
        default: return ew_generic<IONP_FLOAT64>;
    }
}

ionp_status launch_generic(EwArgs A, const ionp_config *cfg, hipStream_t s) {
    if (A.total == 0) return IONP_OK;
    const int block = (cfg && cfg->block > 0) ? cfg->block : 256;
    // split along the outermost coalesced dim so each launch has < 2^31 elements
    const int last = A.ndim - 1;
    int64_t inner = 1;
    for (int d = 0; d < last; d++) inner *= A.shape[d];
    if (inner >= (1ll << 31)) return set_error(IONP_ERR_UNSUPPORTED, "inner extent too large");
    for (int d = 0; d < A.ndim; d++) A.div[d] = FastDiv((uint32_t)(A.shape[d] > 0 ? A.shape[d] : 1));
    bool half = A.out_dt == IONP_FLOAT16;
    for (int k = 0; k < A.nin; k++) half = half || A.in_dt[k] == IONP_FLOAT16;
    const int64_t outer = A.shape[last];
    int64_t step = ((1ll << 31) - 1) / inner;
    if (step < 1) step = 1;
    for (int64_t o0 = 0; o0 < outer; o0 += step) {
        EwArgs B = A;
        const int64_t n = outer - o0 < step ? outer - o0 : step;
        B.shape[last] = n;
        B.total = inner * n;
        for (int k = 0; k < A.nin; k++) B.in[k] = A.in[k] + o0 * A.in_st[k][last] * itemsize(A.in_dt[k]);
        B.out = A.out + o0 * A.out_st[last] * itemsize(A.out_dt);
        GenericFn fn = generic_kernel(A.loop_dt, A.fpe != nullptr);
        if (half) fn = A.fpe ? ew_half<true> : ew_half<false>;
        hipLaunchKernelGGL(fn, dim3(grid_for(B.total, block, cfg)), dim3(block), 0, s, B);
    }
    return hipGetLastError() == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "kernel launch failed");
}

// ---- rows fast path --------------------------------------------------------
typedef void (*RowFn)(const RowArgs);

template <typename T, int VEC>
static RowFn row_kernel_v(int op, int unroll) {
#define IONP_ROW_CASE(OPK)                                                     \
    case OPK:                                                                  \
        return unroll >= 4 ? ew_rows<T, OPK, VEC, 4> : ew_rows<T, OPK, VEC, 1>;
    switch (op) {
        IONP_ROW_CASE(IONP_ADD)
        IONP_ROW_CASE(IONP_SUBTRACT)
        IONP_ROW_CASE(IONP_MULTIPLY)
        IONP_ROW_CASE(IONP_MAXIMUM)
        IONP_ROW_CASE(IONP_MINIMUM)
        IONP_ROW_CASE(IONP_LESS)
        IONP_ROW_CASE(IONP_LESS_EQUAL)
        IONP_ROW_CASE(IONP_GREATER)
        IONP_ROW_CASE(IONP_GREATER_EQUAL)
        IONP_ROW_CASE(IONP_EQUAL)
        IONP_ROW_CASE(IONP_NOT_EQUAL)
        default: return nullptr;
    }
#undef IONP_ROW_CASE
}

template <typename T>
static RowFn row_gather_t(int op) {
    constexpr int V = 16 / sizeof(T) > 4 ? 4 : 16 / sizeof(T);
    switch (op) {
        case IONP_ADD: return ew_rows<T, IONP_ADD, V, 1, true>;
        case IONP_SUBTRACT: return ew_rows<T, IONP_SUBTRACT, V, 1, true>;
        case IONP_MULTIPLY: return ew_rows<T, IONP_MULTIPLY, V, 1, true>;
        case IONP_MAXIMUM: return ew_rows<T, IONP_MAXIMUM, V, 1, true>;
        case IONP_MINIMUM: return ew_rows<T, IONP_MINIMUM, V, 1, true>;
        case IONP_LESS: return ew_rows<T, IONP_LESS, V, 1, true>;
        case IONP_TRUE_DIVIDE:
            if constexpr (std::is_floating_point<T>::value) return ew_rows<T, IONP_TRUE_DIVIDE, V, 1, true>;
            else return nullptr;
        default: return nullptr;
    }
}

// IEEE-flag variants: float add / subtract / multiply / true_divide, one
// vector per thread (as the plain kernels run by default)
template <typename T> static RowFn row_fpe_t(int op, bool gather) {
    constexpr int V = 16 / sizeof(T) > 4 ? 4 : 16 / sizeof(T);
    constexpr int VR = 16 / sizeof(T);
    switch (op) {
        case IONP_ADD: return gather ? ew_rows<T, IONP_ADD, V, 1, true, true> : ew_rows<T, IONP_ADD, VR, 1, false, true>;
        case IONP_SUBTRACT:
            return gather ? ew_rows<T, IONP_SUBTRACT, V, 1, true, true> : ew_rows<T, IONP_SUBTRACT, VR, 1, false, true>;
        case IONP_MULTIPLY:
            return gather ? ew_rows<T, IONP_MULTIPLY, V, 1, true, true> : ew_rows<T, IONP_MULTIPLY, VR, 1, false, true>;
        case IONP_TRUE_DIVIDE:
            return gather ? ew_rows<T, IONP_TRUE_DIVIDE, V, 1, true, true>
                          : ew_rows<T, IONP_TRUE_DIVIDE, 4, 1, false, true>;
        default: return nullptr;
    }
}
static RowFn row_fpe(int dt, int op, bool gather) {
    switch (dt) {
        case IONP_FLOAT32: return row_fpe_t<float>(op, gather);
        case IONP_FLOAT64: return row_fpe_t<double>(op, gather);
        default: return nullptr;
    }
}

static RowFn row_gather(int dt, int op) {
    switch (dt) {
        case IONP_INT32: return row_gather_t<int32_t>(op);
        case IONP_INT64: return row_gather_t<int64_t>(op);
        case IONP_FLOAT32: return row_gather_t<float>(op);
        case IONP_FLOAT64: return row_gather_t<double>(op);
        default: return nullptr;
    }
}

template <typename T>
static RowFn row_kernel_t(int op, int vec, int unroll) {
    if (op == IONP_TRUE_DIVIDE && !std::is_floating_point<T>::value) return nullptr;
    if (op == IONP_TRUE_DIVIDE) {
        switch (vec) {
            case 1: return unroll >= 4 ? ew_rows<T, IONP_TRUE_DIVIDE, 1, 4> : ew_rows<T, IONP_TRUE_DIVIDE, 1, 1>;
            case 2: return unroll >= 4 ? ew_rows<T, IONP_TRUE_DIVIDE, 2, 4> : ew_rows<T, IONP_TRUE_DIVIDE, 2, 1>;
            default: return unroll >= 4 ? ew_rows<T, IONP_TRUE_DIVIDE, 4, 4> : ew_rows<T, IONP_TRUE_DIVIDE, 4, 1>;
        }
    }
    switch (vec) {
        case 1: return row_kernel_v<T, 1>(op, unroll);
        case 2: return row_kernel_v<T, 2>(op, unroll);
        case 4: return row_kernel_v<T, 4>(op, unroll);
        default: return row_kernel_v<T, 8>(op, unroll);
    }
}

static RowFn row_kernel(int dt, int op, int vec, int unroll) {
    switch (dt) {
        case IONP_INT8: return row_kernel_t<int8_t>(op, vec, unroll);
        case IONP_UINT8: return row_kernel_t<uint8_t>(op, vec, unroll);
        case IONP_INT16: return row_kernel_t<int16_t>(op, vec, unroll);
        case IONP_UINT16: return row_kernel_t<uint16_t>(op, vec, unroll);
        case IONP_INT32: return row_kernel_t<int32_t>(op, vec, unroll);
        case IONP_UINT32: return row_kernel_t<uint32_t>(op, vec, unroll);
        case IONP_INT64: return row_kernel_t<int64_t>(op, vec, unroll);
        case IONP_UINT64: return row_kernel_t<uint64_t>(op, vec, unroll);
        case IONP_FLOAT32: return row_kernel_t<float>(op, vec, unroll);
        case IONP_FLOAT64: return row_kernel_t<double>(op, vec, unroll);
        default: return nullptr;
    }
}

// Try the rows fast path for a binary op; returns false if the layout or
// dtypes do not qualify (the caller then uses the generic kernel).
bool try_rows(const EwArgs &A, const ionp_config *cfg, hipStream_t s, ionp_status *st) {
    if (A.nin != 2 || A.op < 0 || A.op >= IONP_NOPS || A.mixed_su) return false;
    if (A.in_dt[0] != A.loop_dt || A.in_dt[1] != A.loop_dt) return false;
    const int want_out = A.op >= IONP_LESS ? (int)IONP_BOOL : A.loop_dt;
    if (A.out_dt != want_out || A.loop_dt == IONP_BOOL) return false;
    if (A.out_st[0] != 1) return false;
    bool gather = false;
    for (int k = 0; k < 2; k++)
        if (A.in_st[k][0] != 0 && A.in_st[k][0] != 1) gather = true;
    if (A.ndim > 4 || A.shape[0] < 64) return false;
    const int esz = itemsize(A.loop_dt);
    int vec = (cfg && cfg->vec > 0) ? cfg->vec : (16 / esz > 8 ? 8 : 16 / esz);
    if (A.op >= IONP_LESS && vec > 4) vec = 4;
    if (gather) vec = 16 / esz > 4 ? 4 : 16 / esz;
    // vector accesses need every contiguous row base aligned to vec elements
    auto aligned = [&](const char *p, int isz, const int64_t *ost, int64_t inner_st) {
        if (inner_st == 0) return true;
        if (((uintptr_t)p) % (uintptr_t)(vec * isz)) return false;
        for (int d = 1; d < A.ndim; d++)
            if (ost[d] % vec) return false;
        return true;
    };
    int64_t ost0[kMaxKDims], ost1[kMaxKDims];
    for (int d = 0; d < A.ndim; d++) {
        ost0[d] = A.in_st[0][d];
        ost1[d] = A.in_st[1][d];
    }
    const int oes = itemsize(A.out_dt);
    // defaults measured on MI355X over the 8192 x 8192 broadcast workloads:
    // one 16-byte vector per thread and one tile per workgroup (no
    // grid-stride loop; the dispatcher balances the short workgroups over
    // the CUs) streams f64 at 6.5 TB/s, against 6.2 with 32 workgroups per
    // CU looping over tiles and 5.6-6.0 with 2-4 vectors in flight
    int unroll = (cfg && cfg->split > 0) ? cfg->split : 1;
    // IEEE flags tracked: the float arithmetic ops have flag variants (one
    // vector width, one vector per thread); other ops raise no flags
    const bool fpe = A.fpe && is_float_dt(A.loop_dt) && A.op <= IONP_TRUE_DIVIDE;
    RowFn fn;
    if (gather) {
        // only the output is accessed by vector; one vector per thread
        // (8192^2 f64 X.T[::2] * c[::2]: 132 us, 141 us with 4 in flight)
        if (A.shape[0] % vec || !aligned(A.out, oes, A.out_st, 1)) return false;
        unroll = 1;
        fn = fpe ? row_fpe(A.loop_dt, A.op, true) : row_gather(A.loop_dt, A.op);
    } else if (fpe) {
        vec = A.op == IONP_TRUE_DIVIDE ? 4 : 16 / esz;
        if (A.shape[0] % vec || !aligned(A.in[0], esz, ost0, A.in_st[0][0]) ||
            !aligned(A.in[1], esz, ost1, A.in_st[1][0]) || !aligned(A.out, oes, A.out_st, 1))
            return false;
        unroll = 1;
        fn = row_fpe(A.loop_dt, A.op, false);
    } else {
        while (vec > 1 && (A.shape[0] % vec || !aligned(A.in[0], esz, ost0, A.in_st[0][0]) ||
                           !aligned(A.in[1], esz, ost1, A.in_st[1][0]) || !aligned(A.out, oes, A.out_st, 1)))
            vec /= 2;
        fn = row_kernel(A.loop_dt, A.op, vec, unroll);
    }
    const int block = (cfg && cfg->block > 0) ? cfg->block : 256;
    if (!fn) return false;
    RowArgs R;
    std::memset(&R, 0, sizeof R);
    R.inner = A.shape[0];
    R.rows = 1;
    R.nouter = A.ndim - 1;
    for (int d = 1; d < A.ndim; d++) {
        R.oshape[d - 1] = A.shape[d];
        R.odiv[d - 1] = FastDiv((uint32_t)A.shape[d]);
        R.ost[0][d - 1] = A.in_st[0][d];
        R.ost[1][d - 1] = A.in_st[1][d];
        R.ost_out[d - 1] = A.out_st[d];
        R.rows *= A.shape[d];
    }
    R.inner_st[0] = A.in_st[0][0];
    R.inner_st[1] = A.in_st[1][0];
    for (int k = 0; k < 2; k++) {  // stride-2 inputs read as aligned pairs of vectors
        bool ok = A.in_st[k][0] == 2 && ((uintptr_t)A.in[k]) % 16 == 0;
        for (int d = 1; ok && d < A.ndim; d++) ok = (A.in_st[k][d] * esz) % 16 == 0;
        R.pair_ok[k] = ok;
    }
    const int64_t tile = (int64_t)block * vec * (unroll >= 4 ? 4 : 1);
    R.tiles_per_row = (R.inner + tile - 1) / tile;
    if (R.rows * R.tiles_per_row >= (1ll << 31) || R.rows >= (1ll << 31)) return false;
    R.tpr = FastDiv((uint32_t)R.tiles_per_row);
    R.in[0] = A.in[0];
    R.in[1] = A.in[1];
    R.out = A.out;
    R.fpe = A.fpe;
    {  // bytes streamed from the inputs (a broadcast row counts once)
        int64_t in_bytes = 0;
        for (int k = 0; k < 2; k++)
            if (A.in_st[k][0] != 0) in_bytes += R.inner * R.rows * esz;
        R.nt_loads = in_bytes > kInfinityCacheBytes;
    }
    const int64_t ntiles = R.rows * R.tiles_per_row;
    // heuristic: a workgroup per tile; a configured blocks_per_cu caps the
    // grid (the kernel loops over the remaining tiles)
    int64_t grid = (cfg && cfg->blocks_per_cu > 0) ? (int64_t)cfg->blocks_per_cu * num_cus() : ntiles;
    if (grid > ntiles) grid = ntiles;
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(block), 0, s, R);
    *st = hipGetLastError() == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "kernel launch failed");
    return true;
}

// ---------------------------------------------------------------------------
// LDS-staged transpose: one operand (Y) is contiguous along a dim d* other
// than the output's contiguous dim d0 (e.g. A + B.T).  A 64 x 64 tile of Y
// is read along d*, where its elements are adjacent, into LDS, and read back
// along d0; the other operand (X) and the output are accessed along d0.
// Every global access of a wavefront is 64 consecutive elements.
// ---------------------------------------------------------------------------
struct TileArgs {
    int64_t n0, n1;        // extents of d0 and d*
    int64_t x_st0, x_st1;  // X strides along d0 and d*
    int64_t y_st0;         // Y stride along d0 (its stride along d* is 1)
    int64_t o_st1;         // output stride along d* (along d0 it is 1)
    int64_t t0, t1;        // tiles along d0, d*
    FastDiv t0div, t1div;
    int32_t nb;            // remaining ("batch") dims
    int64_t bshape[kMaxKDims];
    FastDiv bdiv[kMaxKDims];
    int64_t bx[kMaxKDims], by[kMaxKDims], bo[kMaxKDims];
    int64_t ntiles;
    const char *x, *y;
    char *out;
    int32_t y_first;       // Y is the op's left operand
    uint32_t *fpe;         // IEEE flag word (ew_tiled_v's flag variant only)
};

constexpr int kTile = 64;

template <typename T, int OPK> __global__ __launch_bounds__(256) void ew_tiled(const TileArgs P) {
    using OT = typename std::conditional<(OPK >= IONP_LESS), uint8_t, T>::type;
    __shared__ T tile[kTile][kTile + 1];  // [d0][d*], padded against bank conflicts
    const T *__restrict__ x = (const T *)P.x;
    const T *__restrict__ y = (const T *)P.y;
    OT *__restrict__ o = (OT *)P.out;
    const int lo = threadIdx.x & 63, hi = threadIdx.x >> 6;  // 4 wavefronts
    for (int64_t t = blockIdx.x; t < P.ntiles; t += gridDim.x) {
        // t = (batch, tile along d*, tile along d0), d0 fastest
        uint32_t rem = (uint32_t)t;
        const uint32_t q0 = P.t0div.div(rem);
        const int64_t i0 = (int64_t)(rem - q0 * (uint32_t)P.t0) * kTile;
        rem = q0;
        const uint32_t q1 = P.t1div.div(rem);
        const int64_t i1 = (int64_t)(rem - q1 * (uint32_t)P.t1) * kTile;
        rem = q1;
        int64_t bx = 0, by = 0, bo = 0;
        for (int d = 0; d < P.nb; d++) {
            uint32_t c;
            if (d + 1 < P.nb) {
                const uint32_t q = P.bdiv[d].div(rem);
                c = rem - q * (uint32_t)P.bshape[d];
                rem = q;
            } else {
                c = rem;
            }
            bx += (int64_t)c * P.bx[d];
            by += (int64_t)c * P.by[d];
            bo += (int64_t)c * P.bo[d];
        }
        const int64_t m0 = min((int64_t)kTile, P.n0 - i0), m1 = min((int64_t)kTile, P.n1 - i1);
        // Y tile: lanes walk d* (adjacent in Y)
#pragma unroll
        for (int k = 0; k < kTile / 4; k++) {
            const int r = hi + 4 * k;  // d0 within the tile
            if (r < m0 && lo < m1) tile[r][lo] = y[by + (i0 + r) * P.y_st0 + (i1 + lo)];
        }
        // X along d0 (issued before the barrier)
        T xv[kTile / 4];
#pragma unroll
        for (int k = 0; k < kTile / 4; k++) {
            const int c = hi + 4 * k;  // d* within the tile
            xv[k] = (lo < m0 && c < m1) ? x[bx + (i0 + lo) * P.x_st0 + (i1 + c) * P.x_st1] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kTile / 4; k++) {
            const int c = hi + 4 * k;
            if (lo < m0 && c < m1) {
                const T yv = tile[lo][c];
                o[bo + (i0 + lo) + (i1 + c) * P.o_st1] = P.y_first ? row_apply<T, OPK>(yv, xv[k]) : row_apply<T, OPK>(xv[k], yv);
            }
        }
        __syncthreads();
    }
}

// The vector form: X contiguous (x_st0 == 1) or broadcast (x_st0 == 0)
// along d0, every operand row 16-byte aligned.  A TJ (d0) x TI (d*) tile:
// Y is read in 16-byte vectors along d* (rows of TI elements) into LDS, X
// and the output in 16-byte vectors along d0 (rows of TJ elements), all
// nontemporal (each byte is touched once).  Wide rows give HBM long bursts
// on all three streams: measured on MI355X (8192 x 8192, out = x + y.T),
// 4-byte elements with 128 x 128 tiles 138-143 us (5.6-5.8 TB/s; 64 (d*)
// x 256 (d0) tiles 144-152 us), 8-byte elements with 32 (d*) x 128 (d0)
// tiles 270-278 us (5.8 TB/s), where the scalar 64 x 64 form takes
// 246 / 389 us.  Partial tiles at the edges take
// an element-by-element path.
// XB: X is broadcast along d0 (a template parameter: a runtime branch
// between the loads splits their batch, measured 10 % slower).  FPE: float
// arithmetic with IEEE flags tracked, each element's flags ORed into the
// flag word once per wavefront (NumPy's default modes track them).
template <typename T, int OPK, int TI, int TJ, bool XB, bool FPE = false>
__global__ __launch_bounds__(256) void ew_tiled_v(const TileArgs P) {
    constexpr int V = 16 / sizeof(T);
    using VT = T __attribute__((ext_vector_type(V)));
    using OT = typename std::conditional<(OPK >= IONP_LESS), uint8_t, T>::type;
    using OV = OT __attribute__((ext_vector_type(V)));
    constexpr int YV = TJ * TI / V / 256, XV = TI * TJ / V / 256;  // vectors per thread
    static_assert(YV >= 1 && XV >= 1 && TJ * TI % (256 * V) == 0, "tile shape");
    __shared__ T tile[TJ][TI + 1];  // [d0][d*], padded
    const T *__restrict__ x = (const T *)P.x;
    const T *__restrict__ y = (const T *)P.y;
    OT *__restrict__ o = (OT *)P.out;
    uint32_t fl = 0;
    for (int64_t t = blockIdx.x; t < P.ntiles; t += gridDim.x) {
        // t = (batch, tile along d*, tile along d0), d0 fastest
        uint32_t rem = (uint32_t)t;
        const uint32_t q0 = P.t0div.div(rem);
        const int64_t j0 = (int64_t)(rem - q0 * (uint32_t)P.t0) * TJ;  // along d0
        rem = q0;
        const uint32_t q1 = P.t1div.div(rem);
        const int64_t i0 = (int64_t)(rem - q1 * (uint32_t)P.t1) * TI;  // along d*
        rem = q1;
        int64_t bx = 0, by = 0, bo = 0;
        for (int d = 0; d < P.nb; d++) {
            uint32_t c;
            if (d + 1 < P.nb) {
                const uint32_t q = P.bdiv[d].div(rem);
                c = rem - q * (uint32_t)P.bshape[d];
                rem = q;
            } else {
                c = rem;
            }
            bx += (int64_t)c * P.bx[d];
            by += (int64_t)c * P.by[d];
            bo += (int64_t)c * P.bo[d];
        }
        const int64_t m0 = min((int64_t)TJ, P.n0 - j0), m1 = min((int64_t)TI, P.n1 - i0);
        if (m0 < TJ || m1 < TI) {  // an edge tile: element by element, straight from memory
            for (int e = threadIdx.x; e < TI * TJ; e += blockDim.x) {
                const int i = e / TJ, j = e % TJ;
                if (i >= m1 || j >= m0) continue;
                const T xv = x[bx + (j0 + j) * P.x_st0 + (i0 + i) * P.x_st1];
                const T yv = y[by + (j0 + j) * P.y_st0 + (i0 + i)];
                const OT rv = P.y_first ? row_apply<T, OPK>(yv, xv) : row_apply<T, OPK>(xv, yv);
                o[bo + (j0 + j) + (i0 + i) * P.o_st1] = rv;
                if constexpr (FPE) fl |= P.y_first ? fpe_arith<T, OPK>(yv, xv, rv) : fpe_arith<T, OPK>(xv, yv, rv);
            }
            continue;  // no LDS use: no barrier needed
        }
        VT yv[YV], xv[XV];
#pragma unroll
        for (int k = 0; k < YV; k++) {
            const int idx = threadIdx.x + 256 * k, j = idx / (TI / V), iv = idx % (TI / V);
            yv[k] = __builtin_nontemporal_load((const VT *)(y + by + (j0 + j) * P.y_st0 + i0 + iv * V));
        }
#pragma unroll
        for (int k = 0; k < XV; k++) {
            const int idx = threadIdx.x + 256 * k, i = idx / (TJ / V), jv = idx % (TJ / V);
            const T *px = x + bx + (i0 + i) * P.x_st1;
            if constexpr (!XB) {
                xv[k] = __builtin_nontemporal_load((const VT *)(px + j0 + jv * V));
            } else {
                const T v = *px;  // broadcast along d0
#pragma unroll
                for (int e = 0; e < V; e++) xv[k][e] = v;
            }
        }
#pragma unroll
        for (int k = 0; k < YV; k++) {
            const int idx = threadIdx.x + 256 * k, j = idx / (TI / V), iv = idx % (TI / V);
#pragma unroll
            for (int e = 0; e < V; e++) tile[j][iv * V + e] = yv[k][e];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < XV; k++) {
            const int idx = threadIdx.x + 256 * k, i = idx / (TJ / V), jv = idx % (TJ / V);
            OV r;
#pragma unroll
            for (int e = 0; e < V; e++) {
                const T a = xv[k][e], b = tile[jv * V + e][i];
                r[e] = P.y_first ? row_apply<T, OPK>(b, a) : row_apply<T, OPK>(a, b);
                if constexpr (FPE) fl |= P.y_first ? fpe_arith<T, OPK>(b, a, r[e]) : fpe_arith<T, OPK>(a, b, r[e]);
            }
            __builtin_nontemporal_store(r, (OV *)(o + bo + (i0 + i) * P.o_st1 + j0 + jv * V));
        }
        __syncthreads();
    }
    if constexpr (FPE) fpe_commit(P.fpe, fl);
}

typedef void (*TileFn)(const TileArgs);

// 4-byte elements: 128 x 128 tiles; 8-byte: 32 (d*) x 128 (d0)
template <typename T> constexpr int tile_ti() { return sizeof(T) == 8 ? 32 : 128; }
constexpr int kTileTJ = 128;

template <typename T, bool XB> static TileFn tiled_v_t(int op) {
    constexpr int TI = tile_ti<T>();
#define IONP_TILEV_CASE(OPK) \
    case OPK: return ew_tiled_v<T, OPK, TI, kTileTJ, XB>;
    switch (op) {
        IONP_TILEV_CASE(IONP_ADD)
        IONP_TILEV_CASE(IONP_SUBTRACT)
        IONP_TILEV_CASE(IONP_MULTIPLY)
        IONP_TILEV_CASE(IONP_MAXIMUM)
        IONP_TILEV_CASE(IONP_MINIMUM)
        IONP_TILEV_CASE(IONP_LESS)
        IONP_TILEV_CASE(IONP_LESS_EQUAL)
        IONP_TILEV_CASE(IONP_GREATER)
        IONP_TILEV_CASE(IONP_GREATER_EQUAL)
        IONP_TILEV_CASE(IONP_EQUAL)
        IONP_TILEV_CASE(IONP_NOT_EQUAL)
        case IONP_TRUE_DIVIDE:
            if constexpr (std::is_floating_point<T>::value) return ew_tiled_v<T, IONP_TRUE_DIVIDE, TI, kTileTJ, XB>;
            else return nullptr;
        default: return nullptr;
    }
#undef IONP_TILEV_CASE
}

// the flag variants: float arithmetic only (the ops that raise flags)
template <typename T, bool XB> static TileFn tiled_v_fpe_t(int op) {
    constexpr int TI = tile_ti<T>();
    switch (op) {
        case IONP_ADD: return ew_tiled_v<T, IONP_ADD, TI, kTileTJ, XB, true>;
        case IONP_SUBTRACT: return ew_tiled_v<T, IONP_SUBTRACT, TI, kTileTJ, XB, true>;
        case IONP_MULTIPLY: return ew_tiled_v<T, IONP_MULTIPLY, TI, kTileTJ, XB, true>;
        case IONP_TRUE_DIVIDE: return ew_tiled_v<T, IONP_TRUE_DIVIDE, TI, kTileTJ, XB, true>;
        default: return nullptr;
    }
}

static TileFn tiled_v_kernel(int dt, int op, bool xb, bool fpe) {
    if (fpe) {
        switch (dt) {
            case IONP_FLOAT32: return xb ? tiled_v_fpe_t<float, true>(op) : tiled_v_fpe_t<float, false>(op);
            case IONP_FLOAT64: return xb ? tiled_v_fpe_t<double, true>(op) : tiled_v_fpe_t<double, false>(op);
            default: return nullptr;
        }
    }
    switch (dt) {
        case IONP_INT32: return xb ? tiled_v_t<int32_t, true>(op) : tiled_v_t<int32_t, false>(op);
        case IONP_INT64: return xb ? tiled_v_t<int64_t, true>(op) : tiled_v_t<int64_t, false>(op);
        case IONP_FLOAT32: return xb ? tiled_v_t<float, true>(op) : tiled_v_t<float, false>(op);
        case IONP_FLOAT64: return xb ? tiled_v_t<double, true>(op) : tiled_v_t<double, false>(op);
        default: return nullptr;
    }
}

template <typename T> static TileFn tiled_t(int op) {
#define IONP_TILE_CASE(OPK) \
    case OPK: return ew_tiled<T, OPK>;
    switch (op) {
        IONP_TILE_CASE(IONP_ADD)
        IONP_TILE_CASE(IONP_SUBTRACT)
        IONP_TILE_CASE(IONP_MULTIPLY)
        IONP_TILE_CASE(IONP_MAXIMUM)
        IONP_TILE_CASE(IONP_MINIMUM)
        IONP_TILE_CASE(IONP_LESS)
        IONP_TILE_CASE(IONP_LESS_EQUAL)
        IONP_TILE_CASE(IONP_GREATER)
        IONP_TILE_CASE(IONP_GREATER_EQUAL)
        IONP_TILE_CASE(IONP_EQUAL)
        IONP_TILE_CASE(IONP_NOT_EQUAL)
        case IONP_TRUE_DIVIDE:
            if constexpr (std::is_floating_point<T>::value) return ew_tiled<T, IONP_TRUE_DIVIDE>;
            else return nullptr;
        default: return nullptr;
    }
#undef IONP_TILE_CASE
}

static TileFn tiled_kernel(int dt, int op) {
    switch (dt) {
        case IONP_INT32: return tiled_t<int32_t>(op);
        case IONP_INT64: return tiled_t<int64_t>(op);
        case IONP_FLOAT32: return tiled_t<float>(op);
        case IONP_FLOAT64: return tiled_t<double>(op);
        default: return nullptr;
    }
}

// Binary calls where one operand is contiguous along a dim other than the
// output's contiguous dim 0 (the other operand is read along dim 0 at
// whatever stride it has).
bool try_tiled(const EwArgs &A, const ionp_config *cfg, hipStream_t s, ionp_status *st) {
    if (A.nin != 2 || A.op < 0 || A.op >= IONP_NOPS || A.mixed_su || A.ndim < 2) return false;
    // float arithmetic with IEEE flags tracked: the vector form's flag
    // variant, else nothing here (the LDS tile kernel has flags)
    const bool fpe = A.fpe && is_float_dt(A.loop_dt) && A.op <= IONP_TRUE_DIVIDE;
    if (A.in_dt[0] != A.loop_dt || A.in_dt[1] != A.loop_dt) return false;
    if (A.out_dt != (A.op >= IONP_LESS ? (int)IONP_BOOL : A.loop_dt) || A.out_st[0] != 1) return false;
    int yk = -1, dstar = -1;
    for (int k = 0; k < 2 && yk < 0; k++) {
        const int64_t s0 = A.in_st[k][0];
        if (s0 == 0 || s0 == 1 || s0 == -1) continue;
        for (int d = 1; d < A.ndim; d++)
            if (A.in_st[k][d] == 1) {
                yk = k;
                dstar = d;
                break;
            }
    }
    if (yk < 0) return false;
    const int xk = 1 - yk;
    const TileFn fn = tiled_kernel(A.loop_dt, A.op);
    if (!fn) return false;
    TileArgs P;
    std::memset(&P, 0, sizeof P);
    P.n0 = A.shape[0];
    P.n1 = A.shape[dstar];
    P.x_st0 = A.in_st[xk][0];
    P.x_st1 = A.in_st[xk][dstar];
    P.y_st0 = A.in_st[yk][0];
    P.o_st1 = A.out_st[dstar];
    P.t0 = (P.n0 + kTile - 1) / kTile;
    P.t1 = (P.n1 + kTile - 1) / kTile;
    int64_t nbatch = 1;
    for (int d = 1; d < A.ndim; d++) {
        if (d == dstar) continue;
        P.bshape[P.nb] = A.shape[d];
        P.bdiv[P.nb] = FastDiv((uint32_t)A.shape[d]);
        P.bx[P.nb] = A.in_st[xk][d];
        P.by[P.nb] = A.in_st[yk][d];
        P.bo[P.nb] = A.out_st[d];
        P.nb++;
        nbatch *= A.shape[d];
    }
    P.ntiles = P.t0 * P.t1 * nbatch;
    if (P.ntiles >= (1ll << 31) || P.t0 >= (1ll << 31) || P.t1 >= (1ll << 31)) return false;
    P.t0div = FastDiv((uint32_t)P.t0);
    P.t1div = FastDiv((uint32_t)P.t1);
    P.x = A.in[xk];
    P.y = A.in[yk];
    P.out = A.out;
    P.y_first = yk == 0;
    // the vector form: X contiguous or broadcast along d0 and every row (of
    // every batch) 16-byte aligned, output rows aligned to one vector
    const int64_t isz = itemsize(A.loop_dt), V = 16 / isz, osz = itemsize(A.out_dt);
    auto al = [](const void *p, int64_t a) { return ((uintptr_t)p % (uintptr_t)a) == 0; };
    bool vec = (P.x_st0 == 1 || P.x_st0 == 0) && al(P.y, 16) && (P.y_st0 * isz) % 16 == 0 && al(P.out, V * osz) &&
               P.o_st1 % V == 0 && (P.x_st0 == 0 || (al(P.x, 16) && (P.x_st1 * isz) % 16 == 0));
    for (int d = 0; vec && d < P.nb; d++)
        vec = (P.by[d] * isz) % 16 == 0 && P.bo[d] % V == 0 && (P.x_st0 == 0 || (P.bx[d] * isz) % 16 == 0);
    P.fpe = fpe ? A.fpe : nullptr;
    const TileFn vfn = vec ? tiled_v_kernel(A.loop_dt, A.op, P.x_st0 == 0, fpe) : nullptr;
    if (vfn) {
        const int ti = isz == 8 ? 32 : 128;
        P.t0 = (P.n0 + kTileTJ - 1) / kTileTJ;
        P.t1 = (P.n1 + ti - 1) / ti;
        P.ntiles = P.t0 * P.t1 * nbatch;
        if (P.ntiles >= (1ll << 31) || P.t0 >= (1ll << 31) || P.t1 >= (1ll << 31)) return false;
        P.t0div = FastDiv((uint32_t)P.t0);
        P.t1div = FastDiv((uint32_t)P.t1);
        // measured (8192 x 8192): 2 workgroups per CU (4-byte: 66 KB of LDS
        // each, 2 fit), 1-2 for 8-byte
        const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 2;
        int64_t grid = (int64_t)bpc * num_cus();
        if (grid > P.ntiles) grid = P.ntiles;
        hipLaunchKernelGGL(vfn, dim3((unsigned)grid), dim3(256), 0, s, P);
        *st = hipGetLastError() == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "kernel launch failed");
        return true;
    }
    if (fpe) return false;  // the scalar form has no flag variant
    // measured on MI355X (8192 x 8192, A + B.T): 4 workgroups per CU for
    // 8-byte elements, 16 for 4-byte ones (their LDS tile is half the size)
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : (itemsize(A.loop_dt) == 8 ? 4 : 16);
    int64_t grid = (int64_t)bpc * num_cus();
    if (grid > P.ntiles) grid = P.ntiles;
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(256), 0, s, P);
    *st = hipGetLastError() == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "kernel launch failed");
    return true;
}

ionp_status run_elementwise(EwArgs &A, const ionp_config *cfg, hipStream_t s) {
    if (A.total == 0) return IONP_OK;
    bool half = A.out_dt == IONP_FLOAT16;
    for (int k = 0; k < A.nin; k++) half = half || A.in_dt[k] == IONP_FLOAT16;
    if (half) return launch_generic(A, cfg, s);  // float16: the generic walk only
    ionp_status st;
    if (try_tiled(A, cfg, s, &st)) return st;
    // with IEEE flags tracked try_tiled takes float arithmetic over a
    // transposed operand only in its vector form; the LDS tile kernel
    // (typed, with flags) takes the rest before the rows gather, which
    // would read that operand at a stride (8192^2 f64 A + B.T: 550 us in
    // the gather's flag variant, 313.5 us in the tile kernel's)
    if (A.fpe && try_tiled_any(A, cfg, s, &st)) return st;
    if (try_rows(A, cfg, s, &st)) return st;
    if (try_rows_select_cast(A, cfg, s, &st)) return st;
    if (try_tiled_any(A, cfg, s, &st)) return st;
    return launch_generic(A, cfg, s);
}

}  // namespace ionp
