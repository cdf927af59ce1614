// features.cpp -- layout features of a call for the launch-configuration
// cost model.  No counterpart in the reference.  Host only.
#include <math.h>

#include "planner.h"

using namespace ionp;

extern "C" ionp_status ionp_call_features(const ionp_call *call, float *f) {
    if (!call || !f) return set_error(IONP_ERR_VALUE, "null argument");
    for (int i = 0; i < IONP_NFEATURES; i++) f[i] = 0.0f;
    const int nin = call->kind == IONP_CALL_BINARY ? 2 : call->kind == IONP_CALL_WHERE ? 3 : 1;
    ionp_status s;
    for (int k = 0; k < nin; k++)
        if ((s = check_desc(&call->in[k], "input"))) return s;
    if ((s = check_desc(&call->out, "out"))) return s;
    // iteration space: the output for elementwise calls, the input for reductions
    int nd;
    int64_t shape[IONP_MAX_DIMS];
    int64_t st[4][IONP_MAX_DIMS];
    int nop;
    bool wide = false;
    if (call->kind == IONP_CALL_REDUCE) {
        const ionp_array_desc *in = &call->in[0];
        nd = in->ndim;
        std::memcpy(shape, in->shape, sizeof(int64_t) * nd);
        std::memcpy(st[0], in->strides, sizeof(int64_t) * nd);
        nop = 1;
        bool mask[IONP_MAX_DIMS];
        if ((s = normalize_axes(call->axes, call->naxes, nd, mask))) return s;
        int perm[IONP_MAX_DIMS];
        iter_axis_order(nd, shape, 1, st, perm);
        for (int k = 0; k < nd; k++)
            if (shape[perm[k]] > 1) {
                wide = mask[perm[k]];
                break;
            }
    } else {
        nd = call->out.ndim;
        std::memcpy(shape, call->out.shape, sizeof(int64_t) * nd);
        for (int k = 0; k < nin; k++) broadcast_strides(&call->in[k], nd, shape, st[k]);
        nop = nin;
    }
    // coalesce in iterator order (innermost first), as the kernels do
    int perm[IONP_MAX_DIMS];
    iter_axis_order(nd, shape, nop, st, perm);
    int64_t dims[IONP_MAX_DIMS], cst[4][IONP_MAX_DIMS];
    int n = 0;
    for (int k = 0; k < nd; k++) {
        const int ax = perm[k];
        if (shape[ax] == 1) continue;
        bool merge = n > 0;
        for (int o = 0; merge && o < nop; o++)
            if (st[o][ax] != cst[o][n - 1] * dims[n - 1]) merge = false;
        if (merge) {
            dims[n - 1] *= shape[ax];
        } else {
            dims[n] = shape[ax];
            for (int o = 0; o < nop; o++) cst[o][n] = st[o][ax];
            n++;
        }
    }
    int64_t total = 1;
    for (int k = 0; k < n; k++) total *= dims[k];
    f[0] = (float)n / 4.0f;
    for (int k = 0; k < 4 && k < n; k++) f[1 + k] = (float)log2((double)dims[k] + 1.0) / 24.0f;
    f[5] = (float)log2((double)total + 1.0) / 32.0f;
    for (int o = 0; o < nop && o < 3; o++) {
        f[6 + o] = n > 0 && cst[o][0] == 1 ? 1.0f : 0.0f;
        f[9 + o] = n > 0 && cst[o][0] == 0 ? 1.0f : 0.0f;
    }
    f[12] = 1.0f;
    if (call->kind != IONP_CALL_REDUCE && n > 0) {
        int64_t ost[IONP_MAX_DIMS];
        broadcast_strides(&call->out, nd, shape, ost);
        f[12] = ost[perm[0]] == 1 ? 1.0f : 0.0f;
    }
    const int dt = call->kind == IONP_CALL_WHERE ? call->in[1].dtype : call->in[0].dtype;
    f[13 + (dt == IONP_FLOAT32 ? 0 : dt == IONP_FLOAT64 ? 1 : (dt == IONP_INT64 || dt == IONP_UINT64) ? 2 : 3)] = 1.0f;
    int cls;
    switch (call->kind) {
        case IONP_CALL_BINARY: cls = call->op >= IONP_LESS ? 1 : 0; break;
        case IONP_CALL_WHERE: cls = 2; break;
        case IONP_CALL_CAST: cls = 3; break;
        default: cls = wide ? 4 : 5; break;
    }
    f[17 + cls] = 1.0f;
    f[23] = (float)itemsize(dt) / 8.0f;
    f[24] = (float)itemsize(call->out.dtype) / 8.0f;
    int64_t maxst = 0;
    for (int o = 0; o < nop && o < 3 && n > 0; o++) {
        if (cst[o][0] < 0) f[25 + o] = 1.0f;
        const int64_t a = cst[o][0] < 0 ? -cst[o][0] : cst[o][0];
        if (a > maxst) maxst = a;
    }
    f[28] = (float)log2((double)maxst + 1.0) / 24.0f;
    double bytes = (double)total * itemsize(call->out.dtype);
    for (int o = 0; o < nin; o++) {
        int64_t e = 1;  // elements of the operand itself (a broadcast operand counts its own size)
        for (int k = 0; k < call->in[o].ndim; k++) e *= call->in[o].shape[k];
        bytes += (double)e * itemsize(call->in[o].dtype);
    }
    if (call->kind == IONP_CALL_REDUCE) {
        int64_t e = 1;
        for (int k = 0; k < call->out.ndim; k++) e *= call->out.shape[k];
        bytes = (double)total * itemsize(call->in[0].dtype) + (double)e * itemsize(call->out.dtype);
    }
    f[29] = (float)log2(bytes + 1.0) / 36.0f;
    return IONP_OK;
}
