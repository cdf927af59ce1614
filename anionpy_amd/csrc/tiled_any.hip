// tiled_any.hip -- the LDS-staged transpose for any elementwise call.
//
// Replaces the per-element gathers of binary_elementwise (ionp-core/src/
// ufunc.rs:3539), where_select (sort.rs:1885-1913) and Buffer::cast_to
// (buffer.rs:81-210) when an input is contiguous along an axis other than
// the output's contiguous one (B.T against a C-ordered output), for the
// calls the typed tile kernels do not take: where, casts, mixed dtypes,
// and the IEEE-flag variants.  Output dim 0 is contiguous (d0); d* is a
// dim along which some input is contiguous.  Per 64 (d0) x 32 (d*) tile,
// every operand is copied into LDS as raw items, each read along its own
// contiguous dim (lanes along d* for the transposed operands, along d0
// for the others; each thread's 8 loads in flight before its LDS stores),
// then each lane computes the elements of one d0 position from LDS through
// ew_element_src / ew_element_fpe_src -- the generic kernel's own element
// function, so every dtype, conversion, select and comparison rule is
// that kernel's -- and stores them along d0.  (Exact int64 / uint64
// comparisons stay on the generic kernel.)  where selects, casts and
// mixed-dtype arithmetic store a full tile of a 16-byte-aligned output as
// 16-byte vectors instead, W consecutive d0 positions per lane (8192^2
// f64 where(cond, A, B.T) 387 -> 334 us, where(cond.T, A, B) 366 -> 310
// us, f32 where 283 -> 250 us, f64 + f32.T 320 -> 305 us, measured A/B on
// one MI355X).  Measured and not kept: full
// tiles of 16-byte-aligned operands staged as 16-byte vectors (every
// operand's loads in flight before the LDS stores): 8192^2 f64 where and
// f64 + f32 within 5 % either way, f32 where 9 % and unaligned (8190 x
// 8191) calls 18 % slower, the vector path's registers costing occupancy
// on every call; vectors for the transposed or the d0-contiguous operands
// alone were 20-55 % slower.
#include "elementwise.h"

#include "device_ops.h"

namespace ionp {

constexpr int kTn0 = 64;             // tile extent along d0 (the output's contiguous dim)
constexpr int kTn1 = 32;             // tile extent along d*
constexpr int kTnPad = kTn1 + 1;     // LDS row pitch in items (d* fastest)
constexpr int kTnMaxLds = 64 << 10;  // LDS bytes per workgroup at most

struct TileNArgs {
    EwArgs E;                 // op, dtypes, operand and output addresses, flag word
    int64_t n0, n1;           // extents along d0 and d*
    int64_t st0[3], st1[3];   // input strides along d0 and d*
    int64_t o_st1;            // output stride along d* (along d0 it is 1)
    int32_t staged;           // bit k: input k read along d* (contiguous there)
    int32_t ovec;             // the output's full tiles stored as aligned 16-byte vectors
    int32_t lds_off[3];       // byte offset of input k's LDS tile
    int64_t t0, t1;           // tiles along d0, d*
    FastDiv t0div, t1div;
    int32_t nb;               // remaining ("batch") dims
    int64_t bshape[kMaxKDims];
    FastDiv bdiv[kMaxKDims];
    int64_t bin[3][kMaxKDims], bo[kMaxKDims];
    int64_t ntiles;
};

// One operand's tile into LDS ([d0][d*] items, pitch kTnPad): the 8
// elements of each thread loaded into registers first, then stored.
// along_dstar: lanes walk d* (the operand is contiguous there); else d0.
template <typename U>
__device__ inline void stage_tile(char *tile, const char *src, int64_t st0, int64_t st1, int m0, int m1,
                                  bool along_dstar) {
    constexpr int Q = kTn0 * kTn1 / 256;
    U v[Q];
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int r = along_dstar ? (t >> 5) + 8 * q : (t & 63);
        const int c = along_dstar ? (t & 31) : (t >> 6) + 4 * q;
        v[q] = (r < m0 && c < m1) ? *(const U *)(src + ((int64_t)r * st0 + (int64_t)c * st1) * (int64_t)sizeof(U)) : U(0);
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int r = along_dstar ? (t >> 5) + 8 * q : (t & 63);
        const int c = along_dstar ? (t & 31) : (t >> 6) + 4 * q;
        ((U *)tile)[r * kTnPad + c] = v[q];
    }
}

template <int N> struct UintOf;
template <> struct UintOf<1> { using T = uint8_t; };
template <> struct UintOf<2> { using T = uint16_t; };
template <> struct UintOf<4> { using T = uint32_t; };
template <> struct UintOf<8> { using T = uint64_t; };

// WISZ 0: any call, each element through ew_element_src / _fpe_src in
// loop dtype LT.  WISZ 1 / 2 / 4 / 8: where with a bool condition and
// values of the output's dtype (not bool), a select of WISZ-byte bit
// patterns.  CI >= 0: a cast from dtype CI to CO, the conversion of the
// generic kernel (cvt; a bool input read as 0 / 1) and, FPE, its flags.
// BA >= 0: a binary op on inputs of dtypes BA and BB, both converted to
// the loop dtype LT by cvt, the op applied as the generic kernel applies
// it (output LT, or bool for comparisons) and, FPE, its flags.
template <int LT, bool FPE, int WISZ, int CI = -1, int CO = -1, int BA = -1, int BB = -1>
__global__ __launch_bounds__(256) void ew_tiled_n(const TileNArgs P) {
    extern __shared__ char smem[];
    const EwArgs &A = P.E;
    const int nin = A.nin;
    const int io = itemsize(A.out_dt);
    const int lo = threadIdx.x & 63, hi = threadIdx.x >> 6;  // 4 wavefronts
    uint32_t fl = 0;
    for (int64_t t = blockIdx.x; t < P.ntiles; t += gridDim.x) {
        // t = (batch, tile along d*, tile along d0), d0 fastest
        uint32_t rem = (uint32_t)t;
        const uint32_t q0 = P.t0div.div(rem);
        const int64_t j0 = (int64_t)(rem - q0 * (uint32_t)P.t0) * kTn0;  // along d0
        rem = q0;
        const uint32_t q1 = P.t1div.div(rem);
        const int64_t i0 = (int64_t)(rem - q1 * (uint32_t)P.t1) * kTn1;  // along d*
        rem = q1;
        int64_t b[3] = {0, 0, 0}, bo = 0;
        for (int d = 0; d < P.nb; d++) {
            uint32_t c;
            if (d + 1 < P.nb) {
                const uint32_t q = P.bdiv[d].div(rem);
                c = rem - q * (uint32_t)P.bshape[d];
                rem = q;
            } else {
                c = rem;
            }
            for (int k = 0; k < 3; k++) b[k] += (int64_t)c * P.bin[k][d];
            bo += (int64_t)c * P.bo[d];
        }
        const int m0 = (int)min((int64_t)kTn0, P.n0 - j0), m1 = (int)min((int64_t)kTn1, P.n1 - i0);
        // every operand's tile into LDS, each read along its contiguous dim
        for (int k = 0; k < nin; k++) {
            const int isz = itemsize(A.in_dt[k]);
            char *tile = smem + P.lds_off[k];
            const char *src = A.in[k] + (b[k] + j0 * P.st0[k] + i0 * P.st1[k]) * isz;
            const bool ds = (P.staged >> k) & 1;
            switch (isz) {
                case 1: stage_tile<uint8_t>(tile, src, P.st0[k], P.st1[k], m0, m1, ds); break;
                case 2: stage_tile<uint16_t>(tile, src, P.st0[k], P.st1[k], m0, m1, ds); break;
                case 4: stage_tile<uint32_t>(tile, src, P.st0[k], P.st1[k], m0, m1, ds); break;
                default: stage_tile<uint64_t>(tile, src, P.st0[k], P.st1[k], m0, m1, ds); break;
            }
        }
        __syncthreads();
        if constexpr (WISZ > 0 || CI >= 0 || BA >= 0) {
            // where selects, casts and mixed-dtype arithmetic, a full tile
            // of a 16-byte-aligned output: each lane W consecutive d0
            // positions of one d* position, stored as one nontemporal
            // 16-byte vector (comparisons keep the item stores below)
            using OT = typename std::conditional<
                (WISZ > 0), typename UintOf<(WISZ > 0 ? WISZ : 1)>::T,
                typename std::conditional<(CI >= 0), typename dt_type<(CI >= 0 ? CO : IONP_UINT8)>::T,
                                          typename dt_type<LT>::T>::type>::type;
            if (P.ovec && m0 == kTn0 && m1 == kTn1 && (BA < 0 || A.op < IONP_LESS)) {
                constexpr int W = 16 / (int)sizeof(OT);
                constexpr int NV = kTn0 * kTn1 / W;  // vectors per tile, a multiple of 64
                using OV = OT __attribute__((ext_vector_type(W)));
#pragma unroll
                for (int q = 0; q < (NV + 255) / 256; q++) {
                    const int vi = (int)threadIdx.x + 256 * q;
                    if (NV % 256 != 0 && vi >= NV) break;  // whole wavefronts
                    const int c = vi / (kTn0 / W), r = vi % (kTn0 / W) * W;
                    OV res;
#pragma unroll
                    for (int w = 0; w < W; w++) {
                        const int e = (r + w) * kTnPad + c;
                        if constexpr (WISZ > 0) {
                            const bool cond = ((const uint8_t *)(smem + P.lds_off[0]))[e] != 0;
                            res[w] = cond ? ((const OT *)(smem + P.lds_off[1]))[e] : ((const OT *)(smem + P.lds_off[2]))[e];
                        } else if constexpr (CI >= 0) {
                            using TI = typename dt_type<CI>::T;
                            const TI v = ((const TI *)(smem + P.lds_off[0]))[e];
                            if constexpr (CI == IONP_BOOL) res[w] = cvt<CO>((uint8_t)(v != 0));
                            else res[w] = cvt<CO>(v);
                            if constexpr (FPE && (CI == IONP_FLOAT32 || CI == IONP_FLOAT64)) fl |= fpe_cast_x(CI, CO, (double)v);
                        } else {
                            const OT a = cvt<LT>(((const typename dt_type<(BA >= 0 ? BA : LT)>::T *)(smem + P.lds_off[0]))[e]);
                            const OT bv = cvt<LT>(((const typename dt_type<(BB >= 0 ? BB : LT)>::T *)(smem + P.lds_off[1]))[e]);
                            const OT r = arith_op<LT>(A.op, a, bv);
                            res[w] = r;
                            if constexpr (FPE && std::is_floating_point<OT>::value)
                                if (A.op <= IONP_TRUE_DIVIDE) fl |= fpe_arith_dyn<OT>(A.op, a, bv, r);
                        }
                    }
                    __builtin_nontemporal_store(res, (OV *)(A.out + (bo + j0 + r + (i0 + c) * P.o_st1) * (int64_t)sizeof(OT)));
                }
                __syncthreads();
                continue;
            }
        }
        // lanes along d0 (the output's contiguous dim), 8 positions of d*
        // each, operands read back from LDS
        if (lo < m0) {
            char *orow = A.out + (bo + j0 + lo + i0 * P.o_st1) * io;
            for (int q = 0; q < kTn1 / 4; q++) {
                const int c = hi + 4 * q;
                if (c >= m1) break;
                char *po = orow + (int64_t)c * P.o_st1 * io;
                if constexpr (BA >= 0) {
                    using T = typename dt_type<LT>::T;
                    const int e = lo * kTnPad + c;
                    const T a = cvt<LT>(((const typename dt_type<BA>::T *)(smem + P.lds_off[0]))[e]);
                    const T bv = cvt<LT>(((const typename dt_type<BB>::T *)(smem + P.lds_off[1]))[e]);
                    if (A.op >= IONP_LESS) {
                        *(uint8_t *)po = (uint8_t)cmp_op<T>(A.op, a, bv);
                    } else {
                        const T r = arith_op<LT>(A.op, a, bv);
                        *(T *)po = r;
                        if constexpr (FPE && std::is_floating_point<T>::value)
                            if (A.op <= IONP_TRUE_DIVIDE) fl |= fpe_arith_dyn<T>(A.op, a, bv, r);
                    }
                } else if constexpr (CI >= 0) {
                    using TI = typename dt_type<CI>::T;
                    using TO = typename dt_type<CO>::T;
                    const TI v = ((const TI *)(smem + P.lds_off[0]))[lo * kTnPad + c];
                    if constexpr (CI == IONP_BOOL) *(TO *)po = cvt<CO>((uint8_t)(v != 0));
                    else *(TO *)po = cvt<CO>(v);
                    if constexpr (FPE && (CI == IONP_FLOAT32 || CI == IONP_FLOAT64)) fl |= fpe_cast_x(CI, CO, (double)v);
                } else if constexpr (WISZ > 0) {
                    using U = typename UintOf<WISZ>::T;
                    const int e = lo * kTnPad + c;
                    const bool cond = ((const uint8_t *)(smem + P.lds_off[0]))[e] != 0;
                    *(U *)po = cond ? ((const U *)(smem + P.lds_off[1]))[e] : ((const U *)(smem + P.lds_off[2]))[e];
                } else {
                    PtrSrc src;
                    for (int k = 0; k < 3; k++)
                        src.p[k] = k < nin ? smem + P.lds_off[k] + (lo * kTnPad + c) * itemsize(A.in_dt[k]) : nullptr;
                    if constexpr (FPE) fl |= ew_element_fpe_src<LT>(A, src, po);
                    else ew_element_src<LT>(A, src, po);
                }
            }
        }
        __syncthreads();
    }
    if constexpr (FPE) fpe_commit(A.fpe, fl);
}

typedef void (*TileNFn)(const TileNArgs);

template <bool FPE> static TileNFn tiled_n_kernel(int lt) {
    switch (lt) {
        case IONP_BOOL: return ew_tiled_n<IONP_BOOL, FPE, 0>;
        case IONP_INT8: return ew_tiled_n<IONP_INT8, FPE, 0>;
        case IONP_UINT8: return ew_tiled_n<IONP_UINT8, FPE, 0>;
        case IONP_INT16: return ew_tiled_n<IONP_INT16, FPE, 0>;
        case IONP_UINT16: return ew_tiled_n<IONP_UINT16, FPE, 0>;
        case IONP_INT32: return ew_tiled_n<IONP_INT32, FPE, 0>;
        case IONP_UINT32: return ew_tiled_n<IONP_UINT32, FPE, 0>;
        case IONP_INT64: return ew_tiled_n<IONP_INT64, FPE, 0>;
        case IONP_UINT64: return ew_tiled_n<IONP_UINT64, FPE, 0>;
        case IONP_FLOAT32: return ew_tiled_n<IONP_FLOAT32, FPE, 0>;
        case IONP_FLOAT64: return ew_tiled_n<IONP_FLOAT64, FPE, 0>;
        default: return nullptr;
    }
}

// casts: the flag variant only for float inputs (the others raise none)
template <int CI, bool FPE> static TileNFn tiled_cast_from(int co) {
    switch (co) {
        case IONP_BOOL: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_BOOL>;
        case IONP_INT8: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_INT8>;
        case IONP_UINT8: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_UINT8>;
        case IONP_INT16: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_INT16>;
        case IONP_UINT16: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_UINT16>;
        case IONP_INT32: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_INT32>;
        case IONP_UINT32: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_UINT32>;
        case IONP_INT64: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_INT64>;
        case IONP_UINT64: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_UINT64>;
        case IONP_FLOAT32: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_FLOAT32>;
        case IONP_FLOAT64: return ew_tiled_n<IONP_UINT8, FPE, 0, CI, IONP_FLOAT64>;
        default: return nullptr;
    }
}

static TileNFn tiled_cast_kernel(int ci, int co, bool fpe) {
    switch (ci) {
        case IONP_BOOL: return tiled_cast_from<IONP_BOOL, false>(co);
        case IONP_INT8: return tiled_cast_from<IONP_INT8, false>(co);
        case IONP_UINT8: return tiled_cast_from<IONP_UINT8, false>(co);
        case IONP_INT16: return tiled_cast_from<IONP_INT16, false>(co);
        case IONP_UINT16: return tiled_cast_from<IONP_UINT16, false>(co);
        case IONP_INT32: return tiled_cast_from<IONP_INT32, false>(co);
        case IONP_UINT32: return tiled_cast_from<IONP_UINT32, false>(co);
        case IONP_INT64: return tiled_cast_from<IONP_INT64, false>(co);
        case IONP_UINT64: return tiled_cast_from<IONP_UINT64, false>(co);
        case IONP_FLOAT32: return fpe ? tiled_cast_from<IONP_FLOAT32, true>(co) : tiled_cast_from<IONP_FLOAT32, false>(co);
        case IONP_FLOAT64: return fpe ? tiled_cast_from<IONP_FLOAT64, true>(co) : tiled_cast_from<IONP_FLOAT64, false>(co);
        default: return nullptr;
    }
}

// binaries over int32 / int64 / float32 / float64 (loop dtype: the dtype
// of a same-dtype pair; int64 for two integers, float64 otherwise, as
// NumPy promotes these pairs).  Same-dtype pairs come here when ew_tiled_v
// does not take them (with IEEE flags tracked: rows not 16-byte aligned).
template <int BA, int BB, bool FPE> static TileNFn tiled_bin_pair() {
    constexpr bool ints = (BA == IONP_INT32 || BA == IONP_INT64) && (BB == IONP_INT32 || BB == IONP_INT64);
    return ew_tiled_n<BA == BB ? BA : ints ? IONP_INT64 : IONP_FLOAT64, FPE, 0, -1, -1, BA, BB>;
}
template <int BA, bool FPE> static TileNFn tiled_bin_from(int bb) {
    switch (bb) {
        case IONP_INT32: return tiled_bin_pair<BA, IONP_INT32, FPE>();
        case IONP_INT64: return tiled_bin_pair<BA, IONP_INT64, FPE>();
        case IONP_FLOAT32: return tiled_bin_pair<BA, IONP_FLOAT32, FPE>();
        case IONP_FLOAT64: return tiled_bin_pair<BA, IONP_FLOAT64, FPE>();
        default: return nullptr;
    }
}
template <bool FPE> static TileNFn tiled_bin_kernel(int ba, int bb) {
    switch (ba) {
        case IONP_INT32: return tiled_bin_from<IONP_INT32, FPE>(bb);
        case IONP_INT64: return tiled_bin_from<IONP_INT64, FPE>(bb);
        case IONP_FLOAT32: return tiled_bin_from<IONP_FLOAT32, FPE>(bb);
        case IONP_FLOAT64: return tiled_bin_from<IONP_FLOAT64, FPE>(bb);
        default: return nullptr;
    }
}

static TileNFn tiled_where_kernel(int isz) {
    switch (isz) {
        case 1: return ew_tiled_n<IONP_UINT8, false, 1>;
        case 2: return ew_tiled_n<IONP_UINT8, false, 2>;
        case 4: return ew_tiled_n<IONP_UINT8, false, 4>;
        case 8: return ew_tiled_n<IONP_UINT8, false, 8>;
        default: return nullptr;
    }
}

bool try_tiled_any(const EwArgs &A, const ionp_config *cfg, hipStream_t s, ionp_status *st) {
    if (A.ndim < 2 || A.ndim > kMaxKDims || A.out_st[0] != 1 || A.nin < 1 || A.nin > 3) return false;
    if (A.mixed_su) return false;  // exact int64 / uint64 comparisons: the generic kernel
    if (A.out_dt == IONP_FLOAT16) return false;
    for (int k = 0; k < A.nin; k++)
        if (A.in_dt[k] == IONP_FLOAT16) return false;
    // d*: the first dim (other than 0) along which some input is
    // contiguous while not contiguous or broadcast along d0
    int dstar = -1;
    for (int k = 0; k < A.nin && dstar < 0; k++) {
        const int64_t s0 = A.in_st[k][0];
        if (s0 == 0 || s0 == 1 || s0 == -1) continue;
        for (int d = 1; d < A.ndim; d++)
            if (A.in_st[k][d] == 1) {
                dstar = d;
                break;
            }
    }
    if (dstar < 0) return false;
    // tiles at least half full along d0 and a few elements deep along d*:
    // measured on MI355X, a short output-contiguous dim (2-5 elements, the
    // tile 3-8 % full) ran 7-30 x slower here than in the generic kernel,
    // whose lanes then alternate between a few contiguous rows; a long d0
    // with d* of 6 ran 3 x faster here
    if (A.shape[0] < kTn0 / 2 || A.shape[dstar] < 4) return false;
    // where with a value broadcast along both tile dims (a scalar): the
    // tuning sweep measured the generic kernel 15 % faster over such calls
    // (it streams the two full operands); with both values full or the
    // condition transposed this kernel wins (8192^2 f64: 459 -> 391 us,
    // 620 -> 340 us)
    if (A.op == OP_WHERE)
        for (int k = 1; k < 3; k++)
            if (A.in_st[k][0] == 0 && A.in_st[k][dstar] == 0) return false;
    const bool fpe = A.fpe != nullptr;
    TileNFn fn = fpe ? tiled_n_kernel<true>(A.loop_dt) : tiled_n_kernel<false>(A.loop_dt);
    // where over values of the output's dtype: a bit-pattern select (bool
    // values are read as 0 / 1 by the element function: not a bit copy)
    if (A.op == OP_WHERE && A.in_dt[0] == IONP_BOOL && A.in_dt[1] == A.out_dt && A.in_dt[2] == A.out_dt &&
        A.out_dt != IONP_BOOL)
        fn = tiled_where_kernel(itemsize(A.out_dt));
    // casts: typed conversion (as ew_rows_cast), flags as the generic kernel
    if (A.op == OP_CAST) {
        if (const TileNFn cf = tiled_cast_kernel(A.in_dt[0], A.out_dt, fpe)) fn = cf;
    }
    // binaries: typed loads, the loop dtype NumPy promotes to
    if (A.op >= 0 && A.op < IONP_NOPS && A.nin == 2) {
        const bool ints = (A.in_dt[0] == IONP_INT32 || A.in_dt[0] == IONP_INT64) &&
                          (A.in_dt[1] == IONP_INT32 || A.in_dt[1] == IONP_INT64);
        const int lt = A.in_dt[0] == A.in_dt[1] ? A.in_dt[0] : ints ? IONP_INT64 : IONP_FLOAT64;
        const int odt = A.op >= IONP_LESS ? (int)IONP_BOOL : lt;
        if (A.loop_dt == lt && A.out_dt == odt)
            if (const TileNFn bf = fpe ? tiled_bin_kernel<true>(A.in_dt[0], A.in_dt[1])
                                       : tiled_bin_kernel<false>(A.in_dt[0], A.in_dt[1]))
                fn = bf;
    }
    if (!fn) return false;
    TileNArgs P;
    std::memset(&P, 0, sizeof P);
    P.E = A;
    P.n0 = A.shape[0];
    P.n1 = A.shape[dstar];
    int lds = 0;
    for (int k = 0; k < A.nin; k++) {
        P.st0[k] = A.in_st[k][0];
        P.st1[k] = A.in_st[k][dstar];
        const int64_t s0 = P.st0[k];
        if (P.st1[k] == 1 && s0 != 0 && s0 != 1 && s0 != -1) P.staged |= 1 << k;
        P.lds_off[k] = lds;
        lds += kTn0 * kTnPad * itemsize(A.in_dt[k]);
        lds = (lds + 15) & ~15;
    }
    if (lds > kTnMaxLds) return false;
    P.o_st1 = A.out_st[dstar];
    P.t0 = (P.n0 + kTn0 - 1) / kTn0;
    P.t1 = (P.n1 + kTn1 - 1) / kTn1;
    int64_t nbatch = 1;
    for (int d = 1; d < A.ndim; d++) {
        if (d == dstar) continue;
        P.bshape[P.nb] = A.shape[d];
        P.bdiv[P.nb] = FastDiv((uint32_t)A.shape[d]);
        for (int k = 0; k < A.nin; k++) P.bin[k][P.nb] = A.in_st[k][d];
        P.bo[P.nb] = A.out_st[d];
        P.nb++;
        nbatch *= A.shape[d];
    }
    // 16-byte output vectors: base, stride along d* and batch strides
    // multiples of 16 bytes (tile origins advance by 64 items along d0)
    {
        const int64_t io = itemsize(A.out_dt);
        bool v = reinterpret_cast<uintptr_t>(A.out) % 16 == 0 && (P.o_st1 * io) % 16 == 0;
        for (int d = 0; d < P.nb; d++) v = v && (P.bo[d] * io) % 16 == 0;
        P.ovec = v;
    }
    P.ntiles = P.t0 * P.t1 * nbatch;
    if (P.ntiles >= (1ll << 31) || P.t0 >= (1ll << 31) || P.t1 >= (1ll << 31)) return false;
    P.t0div = FastDiv((uint32_t)P.t0);
    P.t1div = FastDiv((uint32_t)P.t1);
    // a workgroup per tile; a configured blocks_per_cu caps the grid (the
    // kernel loops over the remaining tiles)
    int64_t grid = (cfg && cfg->blocks_per_cu > 0) ? (int64_t)cfg->blocks_per_cu * num_cus() : P.ntiles;
    if (grid > P.ntiles) grid = P.ntiles;
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(256), lds, s, P);
    *st = hipGetLastError() == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, "kernel launch failed");
    return true;
}

}  // namespace ionp
