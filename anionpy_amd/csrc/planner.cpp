// planner.cpp -- host-side planning (no device work); see planner.h.
#include "planner.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

namespace ionp {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_error;

ionp_status set_error(ionp_status code, const std::string &msg) {
    g_error = msg;
    return code;
}
const char *last_error() { return g_error.c_str(); }

// error.rs:54-62: NumPy's bare-comma shape formatting, "(3,)" / "(4,5)"
std::string fmt_shape(const int64_t *shape, int n) {
    std::string s = "(";
    for (int i = 0; i < n; i++) {
        if (i) s += ",";
        s += std::to_string((long long)shape[i]);
    }
    if (n == 1) s += ",";
    return s + ")";
}

const char *dtype_name(int dt) {
    static const char *names[] = {"bool", "int8", "uint8", "int16", "uint16", "int32",
                                  "uint32", "int64", "uint64", "float32", "float64", "float16"};
    return (dt >= 0 && dt < IONP_NTYPES) ? names[dt] : "?";
}
const char *op_name(int op) {
    static const char *names[] = {"add", "subtract", "multiply", "divide", "maximum", "minimum",
                                  "less", "less_equal", "greater", "greater_equal", "equal", "not_equal"};
    return (op >= 0 && op < IONP_NOPS) ? names[op] : "?";
}

// ---------------------------------------------------------------------------
// dtype resolution
// ---------------------------------------------------------------------------
// promote_dtype (dtype.rs:297-357) over the numeric subset: the kind order
// bool < unsigned/signed int < float; mixed signedness widens to the next
// signed size that holds both (uint64 with a signed int -> float64); an
// integer of more than 16 bits with float32 -> float64.
int promote(int a, int b) {
    if (a == b) return a;
    if (a == IONP_BOOL) return b;
    if (b == IONP_BOOL) return a;
    if (a == IONP_FLOAT16 || b == IONP_FLOAT16) {  // float16 holds 8-bit integers exactly
        const int o = a == IONP_FLOAT16 ? b : a;
        if (itemsize(o) == 1) return IONP_FLOAT16;
        if (o == IONP_FLOAT32 || itemsize(o) == 2) return IONP_FLOAT32;
        return IONP_FLOAT64;
    }
    const bool fa = is_float_dt(a), fb = is_float_dt(b);
    if (fa || fb) {
        if (fa && fb) return itemsize(a) >= itemsize(b) ? a : b;
        const int f = fa ? a : b, i = fa ? b : a;
        if (f == IONP_FLOAT64) return IONP_FLOAT64;
        return itemsize(i) <= 2 ? IONP_FLOAT32 : IONP_FLOAT64;
    }
    const bool sa = is_signed_dt(a), sb = is_signed_dt(b);
    const int za = itemsize(a), zb = itemsize(b);
    static const int sgn[9] = {0, IONP_INT8, IONP_INT16, 0, IONP_INT32, 0, 0, 0, IONP_INT64};
    static const int uns[9] = {0, IONP_UINT8, IONP_UINT16, 0, IONP_UINT32, 0, 0, 0, IONP_UINT64};
    if (sa == sb) return sa ? sgn[za > zb ? za : zb] : uns[za > zb ? za : zb];
    const int zs = sa ? za : zb, zu = sa ? zb : za;
    if (zu < zs) return sgn[zs];
    if (zu == 8) return IONP_FLOAT64;
    return sgn[zu * 2];
}

ionp_status binary_types(int op, int a, int b, int *loop, int *out) {
    if (op < 0 || op >= IONP_NOPS) return set_error(IONP_ERR_VALUE, "invalid ufunc op code");
    if (a < 0 || a >= IONP_NTYPES || b < 0 || b >= IONP_NTYPES)
        return set_error(IONP_ERR_TYPE, "invalid dtype code");
    int rt = promote(a, b);
    if (op == IONP_TRUE_DIVIDE) {
        if (!is_real_float(rt)) rt = IONP_FLOAT64;
        *loop = *out = rt;
        return IONP_OK;
    }
    if (op == IONP_SUBTRACT && rt == IONP_BOOL)
        return set_error(IONP_ERR_TYPE,
                         "numpy boolean subtract, the `-` operator, is not supported, use the bitwise_xor, "
                         "the `^` operator, or the logical_xor function instead.");
    *loop = rt;
    *out = op >= IONP_LESS ? (int)IONP_BOOL : rt;
    return IONP_OK;
}

// dtype::can_cast (ionp-core/src/dtype.rs) over the numeric subset.
// safe: bool to anything; an integer to a wider one of its signedness, to a
// signed one wider than it, or to the float whose mantissa holds it (8 bits
// float16, 16 bits float32, else float64); a float to a wider float.
// same_kind: safe, or any cast that keeps or raises the kind (bool < unsigned
// < signed < float).  equiv / no: the same dtype (all of these are native
// byte order).  unsafe: anything.
bool can_cast(int from, int to, int casting) {
    if (from == to || casting == IONP_CAST_UNSAFE) return true;
    if (casting == IONP_CAST_NO || casting == IONP_CAST_EQUIV) return false;
    auto kind = [](int d) { return d == IONP_BOOL ? 0 : is_real_float(d) ? 3 : is_signed_dt(d) ? 2 : 1; };
    const int kf = kind(from), kt = kind(to);
    bool safe;
    if (kf == 0) safe = true;
    else if (kf == 3) safe = kt == 3 && itemsize(to) >= itemsize(from);
    else if (kt == 3) safe = itemsize(from) <= 1 ? true : itemsize(from) == 2 ? to != IONP_FLOAT16 : to == IONP_FLOAT64;
    else if (kt == 0) safe = false;
    else if (kf == kt) safe = itemsize(to) >= itemsize(from);
    else safe = kf == 1 && kt == 2 && itemsize(to) > itemsize(from);  // unsigned into a wider signed
    if (safe || casting == IONP_CAST_SAFE) return safe;
    return kt >= kf;  // same_kind
}

// The loop Ufunc.__call__ runs for `dtype=requested` (ionp-py/src/lib.rs:
// 9568-9804 with resolve_binary_output_dtype, ionp-core/src/ufunc.rs:1031):
// among the loops whose output is `requested`, the first (NumPy's table
// order) that both operands reach by safe casting, else the one whose inputs
// are `requested` itself; the operands must then reach its inputs under
// `casting`.  Arithmetic, maximum and minimum have one loop per dtype
// (T, T -> T; subtract none for bool, true_divide only the floats), so the
// loop is `requested`; comparisons (T, T -> bool) run the loop their
// promotion picks, and no comparison loop has another output.
ionp_status resolve_dtype(int op, int a, int b, int requested, int casting, int *loop) {
    if (op < 0 || op >= IONP_NOPS) return set_error(IONP_ERR_VALUE, "invalid ufunc op code");
    if (a < 0 || a >= IONP_NTYPES || b < 0 || b >= IONP_NTYPES || requested < 0 || requested >= IONP_NTYPES)
        return set_error(IONP_ERR_TYPE, "invalid dtype code");
    if (casting < IONP_CAST_NO || casting > IONP_CAST_UNSAFE) return set_error(IONP_ERR_VALUE, "invalid casting rule");
    const std::string no_loop =
        std::string("No loop matching the specified signature and casting was found for ufunc ") + op_name(op);
    if (op >= IONP_LESS) {
        if (requested != IONP_BOOL) return set_error(IONP_ERR_TYPE, no_loop);
        int l, o;
        ionp_status s = binary_types(op, a, b, &l, &o);
        if (s) return s;
        *loop = l;
        return IONP_OK;
    }
    if (op == IONP_SUBTRACT && requested == IONP_BOOL)
        return set_error(IONP_ERR_TYPE,
                         "numpy boolean subtract, the `-` operator, is not supported, use the bitwise_xor, "
                         "the `^` operator, or the logical_xor function instead.");
    if (op == IONP_TRUE_DIVIDE && !is_real_float(requested)) return set_error(IONP_ERR_TYPE, no_loop);
    static const char *rules[] = {"no", "equiv", "safe", "same_kind", "unsafe"};
    const int ins[2] = {a, b};
    for (int i = 0; i < 2; i++)
        if (!can_cast(ins[i], requested, casting))
            return set_error(IONP_ERR_NO_LOOP, std::string("Cannot cast ufunc '") + op_name(op) + "' input " +
                                                   std::to_string(i) + " from dtype('" + dtype_name(ins[i]) +
                                                   "') to dtype('" + dtype_name(requested) + "') with casting rule '" +
                                                   rules[casting] + "'");
    *loop = requested;
    return IONP_OK;
}

// reduce_compute_dtype (ufunc.rs:2041-2066): add.reduce widens bool and
// sub-64-bit integers to the platform long (signedness kept).
int reduce_default_dtype(int op, int in) {
    if (op == IONP_ADD || op == IONP_MULTIPLY) {
        if (in == IONP_BOOL || in == IONP_INT8 || in == IONP_INT16 || in == IONP_INT32) return IONP_INT64;
        if (in == IONP_UINT8 || in == IONP_UINT16 || in == IONP_UINT32) return IONP_UINT64;
    }
    return in;
}

// ---------------------------------------------------------------------------
// descriptors and broadcasting
// ---------------------------------------------------------------------------
ionp_status check_desc(const ionp_array_desc *d, const char *what) {
    if (!d) return set_error(IONP_ERR_VALUE, std::string(what) + ": null descriptor");
    if (d->dtype < 0 || d->dtype >= IONP_NTYPES) return set_error(IONP_ERR_TYPE, std::string(what) + ": bad dtype");
    if (d->ndim < 0 || d->ndim > IONP_MAX_DIMS) return set_error(IONP_ERR_VALUE, std::string(what) + ": bad ndim");
    for (int i = 0; i < d->ndim; i++)
        if (d->shape[i] < 0) return set_error(IONP_ERR_VALUE, std::string(what) + ": negative dimension");
    return IONP_OK;
}

int64_t desc_size(const ionp_array_desc *d) {
    int64_t n = 1;
    for (int i = 0; i < d->ndim; i++) n *= d->shape[i];
    return n;
}

ionp_status broadcast(int n, const ionp_array_desc *const *ds, int *ndim, int64_t *shape) {
    int nd = 0;
    for (int k = 0; k < n; k++) nd = ds[k]->ndim > nd ? ds[k]->ndim : nd;
    for (int i = 0; i < nd; i++) {
        int64_t dim = 1;
        for (int k = 0; k < n; k++) {
            const int off = nd - ds[k]->ndim;
            const int64_t dk = i < off ? 1 : ds[k]->shape[i - off];
            if (dk == 1) continue;
            if (dim != 1 && dim != dk) {
                std::string msg = "operands could not be broadcast together with shapes ";
                for (int j = 0; j < n; j++) msg += fmt_shape(ds[j]->shape, ds[j]->ndim) + " ";
                return set_error(IONP_ERR_BROADCAST, msg);
            }
            dim = dk;
        }
        shape[i] = dim;
    }
    *ndim = nd;
    return IONP_OK;
}

void broadcast_strides(const ionp_array_desc *d, int ndim, const int64_t *shape, int64_t *st) {
    const int off = ndim - d->ndim;
    for (int i = 0; i < ndim; i++) {
        if (i < off) st[i] = 0;
        else st[i] = (d->shape[i - off] == 1 && shape[i] != 1) ? 0 : d->strides[i - off];
    }
}

// npyiter_find_best_axis_ordering, restated: stable insertion sort of the
// axes from reversed C order by |stride|; per operand, a zero stride on
// either axis abstains; an axis pair swaps only if every operand that
// compares agrees (C order wins conflicts).  Size-1 axes compare as 0.
void iter_axis_order(int ndim, const int64_t *shape, int nop, const int64_t (*st)[IONP_MAX_DIMS], int *perm) {
    for (int k = 0; k < ndim; k++) perm[k] = ndim - 1 - k;
    for (int i = 1; i < ndim; i++) {
        const int ax = perm[i];
        int pos = i;
        for (int j = i - 1; j >= 0; j--) {
            const int bx = perm[j];
            bool compared = false, swap = true;
            for (int op = 0; op < nop; op++) {
                const int64_t s_new = shape[ax] == 1 ? 0 : st[op][ax];
                const int64_t s_old = shape[bx] == 1 ? 0 : st[op][bx];
                if (s_new == 0 || s_old == 0) continue;
                if (llabs(s_old) <= llabs(s_new)) swap = false;
                compared = true;
            }
            if (!compared) continue;
            if (!swap) break;
            pos = j;
        }
        if (pos != i) {
            for (int k = i; k > pos; k--) perm[k] = perm[k - 1];
            perm[pos] = ax;
        }
    }
}

void korder_strides(int nop, const int64_t (*st)[IONP_MAX_DIMS], int ndim, const int64_t *shape, int64_t *out) {
    int perm[IONP_MAX_DIMS];
    iter_axis_order(ndim, shape, nop, st, perm);
    int64_t acc = 1;
    for (int k = 0; k < ndim; k++) {
        out[perm[k]] = acc;
        acc *= shape[perm[k]] > 0 ? shape[perm[k]] : 1;
    }
}

// ---------------------------------------------------------------------------
// elementwise plan
// ---------------------------------------------------------------------------
ionp_status plan_elementwise(int op, int nin, const ionp_array_desc *const *ins, const ionp_array_desc *out,
                             int loop_dt, EwArgs *a) {
    const int nd = out->ndim;
    int64_t st[4][IONP_MAX_DIMS];
    memcpy(st[0], out->strides, sizeof(int64_t) * nd);
    for (int k = 0; k < nin; k++) broadcast_strides(ins[k], nd, out->shape, st[k + 1]);
    int perm[IONP_MAX_DIMS];
    iter_axis_order(nd, out->shape, nin + 1, st, perm);  // output first: favour its contiguity

    // coalesce adjacent axes (innermost first), dropping size-1 axes
    int64_t shp[IONP_MAX_DIMS], cst[4][IONP_MAX_DIMS];
    int n = 0;
    for (int k = 0; k < nd; k++) {
        const int ax = perm[k];
        if (out->shape[ax] == 1) continue;
        bool merge = n > 0;
        for (int o = 0; merge && o <= nin; o++)
            if (st[o][ax] != cst[o][n - 1] * shp[n - 1]) merge = false;
        if (merge) {
            shp[n - 1] *= out->shape[ax];
        } else {
            shp[n] = out->shape[ax];
            for (int o = 0; o <= nin; o++) cst[o][n] = st[o][ax];
            n++;
        }
    }
    if (n == 0) {  // 0-d or all size-1
        shp[0] = 1;
        for (int o = 0; o <= nin; o++) cst[o][0] = 0;
        n = 1;
    }
    if (n > kMaxKDims)
        return set_error(IONP_ERR_UNSUPPORTED, "layout needs more than 8 non-mergeable dimensions");
    memset(a, 0, sizeof *a);
    a->op = op;
    a->nin = nin;
    a->ndim = n;
    a->loop_dt = loop_dt;
    a->out_dt = out->dtype;
    a->total = 1;
    for (int i = 0; i < n; i++) {
        a->shape[i] = shp[i];
        a->total *= shp[i];
        a->out_st[i] = cst[0][i];
        for (int k = 0; k < nin; k++) a->in_st[k][i] = cst[k + 1][i];
    }
    for (int k = 0; k < nin; k++) {
        a->in_dt[k] = ins[k]->dtype;
        a->in[k] = elem0(ins[k]);
    }
    a->out = (char *)elem0(out);
    return IONP_OK;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
ionp_status normalize_axes(const int32_t *axes, int32_t naxes, int ndim, bool *mask) {
    for (int i = 0; i < IONP_MAX_DIMS; i++) mask[i] = false;
    if (naxes < 0) {
        for (int i = 0; i < ndim; i++) mask[i] = true;
        return IONP_OK;
    }
    const int eff = ndim > 1 ? ndim : 1;  // 0-d arrays accept axis 0 / -1
    for (int i = 0; i < naxes; i++) {
        const int ax = axes[i], norm = ax < 0 ? ax + eff : ax;
        if (norm < 0 || norm >= eff)
            return set_error(IONP_ERR_AXIS, "axis " + std::to_string(ax) + " is out of bounds for array of dimension " +
                                                std::to_string(ndim));
        if (ndim == 0) continue;
        if (mask[norm]) return set_error(IONP_ERR_VALUE, "duplicate value in 'axis'");
        mask[norm] = true;
    }
    // duplicates on a 0-d array are still duplicates
    if (ndim == 0 && naxes > 1) return set_error(IONP_ERR_VALUE, "duplicate value in 'axis'");
    return IONP_OK;
}

// Iterator plan of reduce_axis (ufunc.rs:7634 -> 7091 -> 6658), following
// the NumPy buffered reduction it must match: axes ordered by the input's
// strides; the innermost run of reduced axes is the wide block W (pairwise
// per buffer window); reduced axes beyond a kept axis fold sequentially.
ionp_status plan_reduce(int op, const ionp_array_desc *in, const bool *axmask, const ionp_array_desc *out,
                        ReducePlan *p, const int64_t *mask_st, const uint8_t *mask) {
    memset(p, 0, sizeof *p);
    p->op = op;
    p->in_dt = in->dtype;
    p->out_dt = out->dtype;
    p->in = elem0(in);
    p->out = (char *)elem0(out);
    p->mask = mask;
    const int nd = in->ndim;
    int64_t st[2][IONP_MAX_DIMS];
    for (int k = 0; k < nd; k++) {
        st[0][k] = in->strides[k];
        st[1][k] = mask ? mask_st[k] : 0;
    }
    const int64_t *mst = st[1];
    int perm[IONP_MAX_DIMS];
    iter_axis_order(nd, in->shape, mask ? 2 : 1, st, perm);

    // output strides of each kept input axis (out may carry keepdims 1s)
    int64_t ost[IONP_MAX_DIMS];
    {
        int j = 0;
        const bool keep = out->ndim == nd;
        for (int k = 0; k < nd; k++) {
            if (axmask[k]) {
                if (keep) j++;
                ost[k] = 0;
            } else {
                ost[k] = out->strides[j++];
            }
        }
    }
    p->wsize = 1;
    p->rsize = 1;
    p->nout = 1;
    bool in_w = true;
    for (int k = 0; k < nd; k++) {
        const int ax = perm[k];
        const int64_t n = in->shape[ax];
        if (axmask[ax]) {
            if (n == 0) { p->wsize = 0; }
            if (n <= 1) continue;
            if (in_w) {
                if (p->nw > 0 && in->strides[ax] == p->w_st[p->nw - 1] * p->w_shape[p->nw - 1] &&
                    mst[ax] == p->w_m_st[p->nw - 1] * p->w_shape[p->nw - 1]) {
                    p->w_shape[p->nw - 1] *= n;
                } else {
                    if (p->nw == kMaxKDims) return set_error(IONP_ERR_UNSUPPORTED, "reduction layout too complex");
                    p->w_shape[p->nw] = n;
                    p->w_m_st[p->nw] = mst[ax];
                    p->w_st[p->nw++] = in->strides[ax];
                }
                if (p->wsize) p->wsize *= n;
            } else {
                if (p->nr > 0 && in->strides[ax] == p->r_st[p->nr - 1] * p->r_shape[p->nr - 1] &&
                    mst[ax] == p->r_m_st[p->nr - 1] * p->r_shape[p->nr - 1]) {
                    p->r_shape[p->nr - 1] *= n;
                } else {
                    if (p->nr == kMaxKDims) return set_error(IONP_ERR_UNSUPPORTED, "reduction layout too complex");
                    p->r_shape[p->nr] = n;
                    p->r_m_st[p->nr] = mst[ax];
                    p->r_st[p->nr++] = in->strides[ax];
                }
                p->rsize *= n;
            }
        } else {
            if (n == 0) p->nout = 0;
            if (n <= 1) continue;
            in_w = false;
            if (p->nk > 0 && in->strides[ax] == p->k_in_st[p->nk - 1] * p->k_shape[p->nk - 1] &&
                ost[ax] == p->k_out_st[p->nk - 1] * p->k_shape[p->nk - 1] &&
                mst[ax] == p->k_m_st[p->nk - 1] * p->k_shape[p->nk - 1]) {
                p->k_shape[p->nk - 1] *= n;
            } else {
                if (p->nk == kMaxKDims) return set_error(IONP_ERR_UNSUPPORTED, "reduction layout too complex");
                p->k_shape[p->nk] = n;
                p->k_in_st[p->nk] = in->strides[ax];
                p->k_m_st[p->nk] = mst[ax];
                p->k_out_st[p->nk++] = ost[ax];
            }
            if (p->nout) p->nout *= n;
        }
    }
    // an empty reduced axis empties the group
    for (int k = 0; k < nd; k++)
        if (axmask[k] && in->shape[k] == 0) p->wsize = 0;
    p->group = p->wsize * p->rsize;
    return IONP_OK;
}

}  // namespace ionp
