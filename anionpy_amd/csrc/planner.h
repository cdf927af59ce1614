// planner.h -- host-side planning for the gfx950 ufunc/reduce backend.
//
// Everything here runs on the host and touches no device memory: dtype
// resolution, broadcasting, NumPy's K-order output layout, axis coalescing
// and the reduction iteration plan.  The kernels consume the plans.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "ionp_common.h"

namespace ionp {

// ---- errors: thread-local status text (IonpError Display, error.rs:64-90)
ionp_status set_error(ionp_status code, const std::string &msg);
const char *last_error();
std::string fmt_shape(const int64_t *shape, int n);
const char *dtype_name(int dt);
const char *op_name(int op);

// ---- dtype resolution (dtype.rs:297-357, ufunc.rs:1101-1196, 2041-2066)
int promote(int a, int b);
ionp_status binary_types(int op, int a, int b, int *loop, int *out);
int reduce_default_dtype(int op, int in);
// dtype::can_cast under an ionp_casting rule
bool can_cast(int from, int to, int casting);
// Ufunc.__call__'s dtype= loop selection (lib.rs:9568-9804, ufunc.rs:1031)
ionp_status resolve_dtype(int op, int a, int b, int requested, int casting, int *loop);

// ---- descriptors
ionp_status check_desc(const ionp_array_desc *d, const char *what);
inline const char *elem0(const ionp_array_desc *d) {
    return (const char *)d->data + d->offset * itemsize(d->dtype);
}
int64_t desc_size(const ionp_array_desc *d);

// Broadcast shapes of n descriptors (shape.rs:54).  On failure records the
// NumPy message naming the shapes.
ionp_status broadcast(int n, const ionp_array_desc *const *ds, int *ndim, int64_t *shape);

// Strides of d viewed at the broadcast shape (shape.rs:113): 0 on broadcast axes.
void broadcast_strides(const ionp_array_desc *d, int ndim, const int64_t *shape, int64_t *st);

// NumPy nditer axis order over `nop` operands (innermost first).
void iter_axis_order(int ndim, const int64_t *shape, int nop, const int64_t (*st)[IONP_MAX_DIMS], int *perm);

// K-order dense strides for an output of `shape` iterated with the operands.
void korder_strides(int nop, const int64_t (*st)[IONP_MAX_DIMS], int ndim, const int64_t *shape, int64_t *out);

// ---- elementwise plan: coalesced iteration space for EwArgs
// Fills shape/strides/dtypes/pointers of `a` from the operands (already
// broadcast-validated) and the output.  Returns IONP_ERR_UNSUPPORTED if the
// coalesced rank exceeds kMaxKDims.
ionp_status plan_elementwise(int op, int nin, const ionp_array_desc *const *ins, const ionp_array_desc *out,
                             int loop_dt, EwArgs *a);

// ---- reduction plan
struct ReducePlan {
    int op;        // IONP_ADD / IONP_MAXIMUM / IONP_MINIMUM
    int in_dt, out_dt;
    // kept axes (coalesced, innermost first) with input and output strides
    int nk;
    int64_t k_shape[kMaxKDims], k_in_st[kMaxKDims], k_out_st[kMaxKDims];
    // wide block W (innermost reduced axes): coalesced runs, innermost first
    int nw;
    int64_t w_shape[kMaxKDims], w_st[kMaxKDims];
    int64_t wsize;
    // outer reduced axes (iterator order, innermost first)
    int nr;
    int64_t r_shape[kMaxKDims], r_st[kMaxKDims];
    int64_t rsize;
    int64_t nout;
    int64_t group;  // elements per output = wsize * rsize
    const char *in;
    char *out;
    // where= mask (bool, broadcast to the input's shape) or null: its
    // strides along the same coalesced kept / W / outer axes
    const uint8_t *mask;
    int64_t k_m_st[kMaxKDims], w_m_st[kMaxKDims], r_m_st[kMaxKDims];
};

// axes: normalized reduced-axis mask; out: output descriptor as planned.
// mask_st / mask: the where= operand's strides broadcast to the input's
// shape and its element-0 address (null: no mask); the mask takes part in
// the axis order and in coalescing, as NumPy's third reduction operand.
ionp_status plan_reduce(int op, const ionp_array_desc *in, const bool *axmask, const ionp_array_desc *out,
                        ReducePlan *p, const int64_t *mask_st = nullptr, const uint8_t *mask = nullptr);

// normalize_reduce_axes (ufunc.rs:7551-7617)
ionp_status normalize_axes(const int32_t *axes, int32_t naxes, int ndim, bool *mask);

}  // namespace ionp
