// reduce.h -- launch interface of the reduction kernels.
#pragma once

#include "planner.h"

namespace ionp {

// Device-side view of a ReducePlan: coalesced kept / wide / outer-reduced
// axes with fast dividers, buffer-window geometry and scratch pointers.
struct RedArgs {
    int32_t in_dt, out_dt, op;
    int32_t nk, nw, nr;
    int64_t nout, wsize, rsize, group;
    int64_t bufsize, nwin;  // add/max windows: nwin = ceil(wsize / bufsize)
    int64_t k_shape[kMaxKDims], k_in_st[kMaxKDims], k_out_st[kMaxKDims];
    FastDiv k_div[kMaxKDims];
    int64_t w_shape[kMaxKDims], w_st[kMaxKDims];
    FastDiv w_div[kMaxKDims];
    int64_t r_shape[kMaxKDims], r_st[kMaxKDims];
    FastDiv r_div[kMaxKDims];
    const char *in;
    char *out;
    void *scratch;  // partial results
    int32_t has_init;    // initial= given: the fold starts from it (and max/min feed every element)
    int32_t pad_;
    uint64_t init_bits;  // its value in the output dtype: float bits, or the integer sign-/zero-extended
    // where= mask (bool bytes) or null, with its strides along the kept / W /
    // outer axes (the dividers above decompose indices for it too)
    const uint8_t *mask;
    int64_t k_m_st[kMaxKDims], w_m_st[kMaxKDims], r_m_st[kMaxKDims];
};

// Float add: every window is one pairwise leaf (<= 128 elements) and there
// are outputs enough (or no outer reduced loop) for one thread per output.
inline bool pw_small_ok(const RedArgs &A) {
    return A.nw > 0 && A.nwin == 1 && A.wsize <= 128 && (A.rsize == 1 || A.nout >= 2048);
}

// Narrow float add (W empty) with long chains and few outputs: a wavefront
// per output keeps enough loads in flight for the serial add chain.
inline bool narrow_wave_ok(const RedArgs &A) { return A.nw == 0 && A.rsize >= 256 && A.nout <= 8192; }

// Integer reductions: one thread per output (elements in any order) when
// groups are short, or outputs plenty and groups moderate.
inline bool int_cols_ok(const RedArgs &A) {
    return A.group <= 64 || (A.nout >= 16384 && A.group <= 4096) || (A.nw == 0 && A.nout >= 2048 && A.group <= 65536);
}

// initial: one element of the output dtype (host memory), or null
ionp_status run_reduce(const ReducePlan &p, int64_t bufsize, const void *initial, const ionp_config *cfg, hipStream_t s);

// RedArgs::init_bits of one output-dtype element at `p`
uint64_t init_bits_of(int out_dt, const void *p);

// argmax / argmin along one axis (C index order) or over the C-order ravel.
struct ArgArgs {
    int32_t in_dt, ismax;
    int64_t nout, len;
    // kept axes (C order of the input without `axis`), with input and output strides
    int32_t nk;
    int64_t k_shape[IONP_MAX_DIMS], k_in_st[IONP_MAX_DIMS], k_out_st[IONP_MAX_DIMS];
    FastDiv k_div[IONP_MAX_DIMS];
    // the scanned index space: one axis (nl == 1) or the C-order ravel
    int32_t nl;
    int64_t l_shape[IONP_MAX_DIMS], l_st[IONP_MAX_DIMS];
    FastDiv l_div[IONP_MAX_DIMS];
    int64_t seg, nseg;
    const char *in;
    char *out;
    void *scratch;
};

ionp_status run_argext(ArgArgs &a, const ionp_config *cfg, hipStream_t s);

// ufunc.reduceat along one axis: output (fiber f, k) reduces the axis
// segment that starts at indices[k] (see run_reduceat).
struct RatArgs {
    int32_t op, in_dt, out_dt, nf;
    int64_t len, in_st, out_st;  // the axis: extent and strides (elements)
    int64_t nfib, nidx;
    int64_t fshape[IONP_MAX_DIMS];  // the other axes, innermost first
    FastDiv fdiv[IONP_MAX_DIMS];
    int64_t fin[IONP_MAX_DIMS], fout[IONP_MAX_DIMS];
    const int64_t *idx;  // device copy of the indices
    const char *in;
    char *out;
};

// indices: host array of nidx entries, each in [0, len) (checked by the caller)
ionp_status run_reduceat(RatArgs &a, const int64_t *indices, const ionp_config *cfg, hipStream_t s);

// device scratch owned by the library (grown on demand, stream-ordered use)
void *scratch_get(size_t bytes, hipStream_t s);
// a second such buffer for inputs converted to another dtype first
void *stage_get(size_t bytes, hipStream_t s);

// Bytes of each scratch kind: reduction partials (scratch_get), converted
// inputs (stage_get) and the in-launch folders' partial slots.
struct ScratchUse {
    size_t scratch = 0, stage = 0, slots = 0;
};

// Scratch regions a caller owns (a batch's captured graph, the
// no-allocation entry point).  While installed on a thread (ScratchScope),
// the getters above return these regions instead of the library's buffers:
// nothing is allocated or freed, the fold slots are refilled with their
// sentinel by a kernel ahead of every use, and a request larger than its
// region returns null and sets `short_` (`need` records the largest
// request of each kind).
struct ScratchArena {
    char *scratch = nullptr, *stage = nullptr, *slots = nullptr;
    ScratchUse size;
    ScratchUse need;
    bool short_ = false;
};
struct ScratchScope {
    explicit ScratchScope(ScratchArena *a);
    ~ScratchScope();
    ScratchArena *old;
};
// True (once) when an in-launch fold of a one-output add.reduce gave up
// waiting for a window partial since the last call; the output it wrote is
// NaN.  Reads host-visible memory: no synchronisation.
bool fold_error_take();
// Test hook: the folders treat window `window` (-1: none) as never filled
// and give up after `polls` polls of one unfilled slot (default 2^26).
void fold_debug_stall(int64_t window, uint32_t polls);

// Records the largest request of each kind made on this thread while alive
// (the library's buffers still serve them).
struct ScratchRecord {
    explicit ScratchRecord(ScratchUse *u);
    ~ScratchRecord();
    ScratchUse *old;
};

}  // namespace ionp
