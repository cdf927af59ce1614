// api.cpp -- the extern "C" entry points declared in include/ionp_hip.h.
//
// Validation happens on the host before any launch, in the order the
// reference validates (dtype loop before shape, ufunc.rs:3629-3640), and
// errors carry NumPy's message text (error.rs:64-90).
#include <atomic>
#include <mutex>
#include <vector>

#include "elementwise.h"
#include "advisor.h"
#include "planner.h"
#include "reduce.h"
#include "accumulate.h"

using namespace ionp;

namespace {

std::atomic<int64_t> g_bufsize{8192};

hipStream_t as_stream(void *s) { return (hipStream_t)s; }

void fill_template(ionp_array_desc *t, int dtype, int ndim, const int64_t *shape, const int64_t *strides) {
    std::memset(t, 0, sizeof *t);
    t->dtype = dtype;
    t->ndim = ndim;
    for (int i = 0; i < ndim; i++) {
        t->shape[i] = shape[i];
        t->strides[i] = strides[i];
    }
}

ionp_status same_shape(const ionp_array_desc *out, int ndim, const int64_t *shape, const char *what) {
    bool ok = out->ndim == ndim;
    for (int i = 0; ok && i < ndim; i++) ok = out->shape[i] == shape[i];
    if (!ok)
        return set_error(IONP_ERR_VALUE, std::string("non-broadcastable output operand with shape ") +
                                             fmt_shape(out->shape, out->ndim) + " doesn't match the broadcast shape " +
                                             fmt_shape(shape, ndim) + " (" + what + ")");
    return IONP_OK;
}

ionp_status check_cfg(const ionp_config *cfg) {
    if (!cfg) return IONP_OK;
    if (cfg->block != 0 && (cfg->block < 64 || cfg->block > 256 || cfg->block % 64))
        return set_error(IONP_ERR_VALUE, "ionp_config.block must be 64, 128, 192 or 256");
    if (cfg->vec != 0 && cfg->vec != 1 && cfg->vec != 2 && cfg->vec != 4 && cfg->vec != 8)
        return set_error(IONP_ERR_VALUE, "ionp_config.vec must be 1, 2, 4 or 8");
    if (cfg->blocks_per_cu < 0 || cfg->blocks_per_cu > 4096 || cfg->split < 0 || cfg->split > 64)
        return set_error(IONP_ERR_VALUE, "ionp_config.blocks_per_cu must be in [0, 4096] and split in [0, 64]");
    return IONP_OK;
}

ionp_status binary_common(int32_t op, const ionp_array_desc *a, const ionp_array_desc *b, int *loop, int *odt,
                          int *nd, int64_t *shape) {
    ionp_status st;
    if ((st = check_desc(a, "x1")) || (st = check_desc(b, "x2"))) return st;
    if ((st = binary_types(op, a->dtype, b->dtype, loop, odt))) return st;
    const ionp_array_desc *ops[2] = {a, b};
    return broadcast(2, ops, nd, shape);
}

// IEEE flag tracking (np.errstate), per thread (fpstate.rs:103 keeps the
// reference's FPE state thread-local): while a thread has it on, its
// ionp_binary and ionp_cast calls OR the flags their elements raise into
// the thread's own word.  The words of one GPU are one array of kFpeSlots;
// a thread takes a slot index on first use and returns it when it ends.
constexpr int kFpeSlots = 1024;
struct FpePool {
    std::mutex mu;
    uint32_t *words[64] = {};  // per device
    uint32_t *host = nullptr;  // pinned, one word per slot: where a fetch copies its word
    std::vector<int> free_slots;
    int next = 0;
};
FpePool &fpe_pool() {
    static FpePool *p = new FpePool();  // never destroyed: threads may end after static destructors ran
    return *p;
}
struct FpeThread {
    int on = 0;
    int slot = -1;
    ~FpeThread() {
        if (slot < 0) return;
        FpePool &P = fpe_pool();
        std::lock_guard<std::mutex> lk(P.mu);
        P.free_slots.push_back(slot);
    }
};
thread_local FpeThread t_fpe;

// this thread's flag word on the current device
ionp_status fpe_word(uint32_t **w) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(IONP_ERR_DEVICE, "no current HIP device");
    FpePool &P = fpe_pool();
    std::lock_guard<std::mutex> lk(P.mu);
    if (t_fpe.slot < 0) {
        if (!P.free_slots.empty()) {
            t_fpe.slot = P.free_slots.back();
            P.free_slots.pop_back();
        } else if (P.next < kFpeSlots) {
            t_fpe.slot = P.next++;
        } else {
            return set_error(IONP_ERR_UNSUPPORTED, "more than 1024 threads track IEEE flags at once");
        }
    }
    if (!P.words[dev]) {
        void *p = nullptr;
        if (hipMalloc(&p, kFpeSlots * sizeof(uint32_t)) != hipSuccess)
            return set_error(IONP_ERR_DEVICE, "out of device memory for the flag words");
        if (hipMemset(p, 0, kFpeSlots * sizeof(uint32_t)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            (void)hipFree(p);
            return set_error(IONP_ERR_DEVICE, "clearing the flag words failed");
        }
        P.words[dev] = (uint32_t *)p;
    }
    if (!P.host) {
        void *h = nullptr;
        if (hipHostMalloc(&h, kFpeSlots * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess)
            return set_error(IONP_ERR_DEVICE, "out of pinned host memory for the flag words");
        P.host = (uint32_t *)h;
    }
    *w = P.words[dev] + t_fpe.slot;
    return IONP_OK;
}

// only loops that can raise a flag get the word (float arithmetic, casts
// from a float or to float16): the others run the plain kernels
ionp_status attach_fpe(EwArgs &A, int op, int loop, int in_dt, int out_dt) {
    if (!t_fpe.on) return IONP_OK;
    const bool can = op == OP_CAST ? is_real_float(in_dt) || out_dt == IONP_FLOAT16
                                   : op <= IONP_TRUE_DIVIDE && is_real_float(loop);
    return can ? fpe_word(&A.fpe) : IONP_OK;
}

// `in` cast to dtype `dt` into the library's stage buffer, dense in the
// input's K order (NumPy's buffered cast ahead of a loop)
ionp_status cast_copy(const ionp_array_desc *in, int dt, ionp_array_desc *c, hipStream_t st) {
    ionp_status s;
    if ((s = ionp_plan_cast(in, dt, c))) return s;
    int64_t nel = 1;
    for (int i = 0; i < in->ndim; i++) nel *= in->shape[i];
    void *tmp = stage_get((size_t)(nel > 0 ? nel : 1) * itemsize(dt), st);
    if (!tmp) return set_error(IONP_ERR_DEVICE, "out of device memory for a converted input");
    c->data = tmp;
    c->offset = 0;
    const ionp_array_desc *ins[1] = {in};
    EwArgs E;
    if ((s = plan_elementwise(OP_CAST, 1, ins, c, dt, &E))) return s;
    return run_elementwise(E, nullptr, st);
}

// A float32 / float64 loop that reads its input through NumPy's buffered
// cast: float16 input (add / multiply convert as they read otherwise), and
// maximum / minimum with another dtype (the x86 loop models hold for the
// loop's own dtype, read from the contiguous cast buffer).  Those loops run
// on a cast copy in the library's stage buffer, dense in the input's K
// order (the iterator's axis order is kept).
bool needs_cast_copy(int op, int in_dt, int out_dt) {
    if (!is_float_dt(out_dt) || in_dt == out_dt) return false;
    return in_dt == IONP_FLOAT16 || op == IONP_MAXIMUM || op == IONP_MINIMUM;
}

bool is_mixed_su(int op, int a, int b) {
    return op >= IONP_LESS && ((a == IONP_INT64 && b == IONP_UINT64) || (a == IONP_UINT64 && b == IONP_INT64));
}

// Scratch of ionp_reduce_noalloc: one fixed array per GPU, in the code
// object, so no call allocates.  Regions: reduction partials, converted
// (float16) inputs, the in-launch folders' slots.
constexpr size_t kNoAllocPartials = 4u << 20, kNoAllocStage = 2u << 20, kNoAllocSlots = 2u << 20;
constexpr size_t kNoAllocBytes = kNoAllocPartials + kNoAllocStage + kNoAllocSlots;
static __device__ __align__(256) char scratch[kNoAllocBytes];

// The array's address on each device (hipGetSymbolAddress with that
// device current; the host-side name `scratch` is not a device address).
std::mutex g_noalloc_mu;
char *g_noalloc_ptr[64] = {};

ionp_status noalloc_base(char **base) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return set_error(IONP_ERR_DEVICE, "no current HIP device");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(IONP_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_noalloc_ptr[dev] = (char *)p;
    }
    *base = g_noalloc_ptr[dev];
    return IONP_OK;
}

}  // namespace

namespace ionp {
int fpe_thread_swap(int on) {
    const int old = t_fpe.on;
    t_fpe.on = on;
    return old;
}
}  // namespace ionp

extern "C" {

ionp_status ionp_init(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return set_error(IONP_ERR_DEVICE, "no HIP device available");
    (void)num_cus();
    char *base;
    return noalloc_base(&base);  // the no-allocation scratch of the current device
}

const char *ionp_last_error(void) { return last_error(); }

ionp_status ionp_fpe_track(int32_t on) {
    // every launch that may still OR into the word finishes, then the word
    // is cleared: flags raised under the old state never reach a later fetch
    uint32_t *w;
    ionp_status s;
    if ((s = fpe_word(&w))) return s;
    if (hipDeviceSynchronize() != hipSuccess || hipMemset(w, 0, sizeof *w) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return set_error(IONP_ERR_DEVICE, "clearing the flag word failed");
    t_fpe.on = on ? 1 : 0;
    return IONP_OK;
}

static ionp_status fold_failed() {
    return set_error(IONP_ERR_DEVICE,
                     "an add.reduce fold gave up waiting for window partials (a producer workgroup was not "
                     "scheduled); that result is NaN");
}

ionp_status ionp_check(void *stream) {
    if (hipStreamSynchronize(as_stream(stream)) != hipSuccess)
        return set_error(IONP_ERR_DEVICE, "hipStreamSynchronize failed");
    return fold_error_take() ? fold_failed() : IONP_OK;
}

ionp_status ionp_debug_fold_stall(int64_t window, uint32_t polls) {
    fold_debug_stall(window, polls);
    return IONP_OK;
}

ionp_status ionp_can_cast(int32_t from, int32_t to, int32_t casting, int32_t *ok) {
    if (from < 0 || from >= IONP_NTYPES || to < 0 || to >= IONP_NTYPES) return set_error(IONP_ERR_TYPE, "invalid dtype code");
    if (casting < IONP_CAST_NO || casting > IONP_CAST_UNSAFE) return set_error(IONP_ERR_VALUE, "invalid casting rule");
    if (!ok) return set_error(IONP_ERR_VALUE, "ok must not be null");
    *ok = can_cast(from, to, casting) ? 1 : 0;
    return IONP_OK;
}

ionp_status ionp_resolve_dtype(int32_t op, int32_t a, int32_t b, int32_t requested, int32_t casting, int32_t *loop) {
    if (!loop) return set_error(IONP_ERR_VALUE, "loop must not be null");
    int l;
    ionp_status s = resolve_dtype(op, a, b, requested, casting, &l);
    if (s) return s;
    *loop = l;
    return IONP_OK;
}

ionp_status ionp_fpe_fetch(void *stream, uint32_t *flags) {
    if (!flags) return set_error(IONP_ERR_VALUE, "flags must not be null");
    uint32_t *w;
    ionp_status s;
    if ((s = fpe_word(&w))) return s;
    // the word is copied behind the stream's launches into this thread's
    // pinned slot, and one synchronisation waits for both
    volatile uint32_t *h = fpe_pool().host + t_fpe.slot;
    if (hipMemcpyAsync((void *)h, w, sizeof *w, hipMemcpyDeviceToHost, as_stream(stream)) != hipSuccess ||
        hipStreamSynchronize(as_stream(stream)) != hipSuccess)
        return set_error(IONP_ERR_DEVICE, "reading the flag word failed");
    const uint32_t v = *h;
    if (v && hipMemset(w, 0, sizeof v) != hipSuccess) return set_error(IONP_ERR_DEVICE, "clearing the flag word failed");
    *flags = v;
    return IONP_OK;
}

int32_t ionp_version(void) { return (1 << 16) | 0; }

ionp_status ionp_set_bufsize(int64_t elements) {
    if (elements < 16 || elements > 8192) return set_error(IONP_ERR_VALUE, "bufsize must be in [16, 8192]");
    g_bufsize = elements;
    return IONP_OK;
}
int64_t ionp_get_bufsize(void) { return g_bufsize; }

ionp_status ionp_result_type(int32_t a, int32_t b, int32_t *out) {
    if (a < 0 || a >= IONP_NTYPES || b < 0 || b >= IONP_NTYPES) return set_error(IONP_ERR_TYPE, "invalid dtype code");
    *out = promote(a, b);
    return IONP_OK;
}

ionp_status ionp_binary_types(int32_t op, int32_t a, int32_t b, int32_t *loop, int32_t *out) {
    int l, o;
    ionp_status st = binary_types(op, a, b, &l, &o);
    if (st) return st;
    *loop = l;
    *out = o;
    return IONP_OK;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
ionp_status ionp_plan_binary(int32_t op, const ionp_array_desc *a, const ionp_array_desc *b,
                             ionp_array_desc *out_template) {
    int loop, odt, nd;
    int64_t shape[IONP_MAX_DIMS], st[2][IONP_MAX_DIMS], ost[IONP_MAX_DIMS];
    ionp_status s = binary_common(op, a, b, &loop, &odt, &nd, shape);
    if (s) return s;
    broadcast_strides(a, nd, shape, st[0]);
    broadcast_strides(b, nd, shape, st[1]);
    korder_strides(2, st, nd, shape, ost);
    fill_template(out_template, odt, nd, shape, ost);
    return IONP_OK;
}

ionp_status ionp_plan_where(const ionp_array_desc *c, const ionp_array_desc *x, const ionp_array_desc *y,
                            ionp_array_desc *out_template) {
    ionp_status s;
    if ((s = check_desc(c, "condition")) || (s = check_desc(x, "x")) || (s = check_desc(y, "y"))) return s;
    const ionp_array_desc *ops[3] = {c, x, y};
    int nd;
    int64_t shape[IONP_MAX_DIMS], st[3][IONP_MAX_DIMS], ost[IONP_MAX_DIMS];
    if ((s = broadcast(3, ops, &nd, shape))) return s;
    for (int k = 0; k < 3; k++) broadcast_strides(ops[k], nd, shape, st[k]);
    korder_strides(3, st, nd, shape, ost);
    fill_template(out_template, promote(x->dtype, y->dtype), nd, shape, ost);
    return IONP_OK;
}

ionp_status ionp_plan_cast(const ionp_array_desc *in, int32_t dtype, ionp_array_desc *out_template) {
    ionp_status s;
    if ((s = check_desc(in, "array"))) return s;
    if (dtype < 0 || dtype >= IONP_NTYPES) return set_error(IONP_ERR_TYPE, "invalid dtype code");
    int64_t st[1][IONP_MAX_DIMS], ost[IONP_MAX_DIMS];
    std::memcpy(st[0], in->strides, sizeof(int64_t) * in->ndim);
    korder_strides(1, st, in->ndim, in->shape, ost);
    fill_template(out_template, dtype, in->ndim, in->shape, ost);
    return IONP_OK;
}

static ionp_status reduce_dtype(int32_t op, int in_dt, int32_t dtype, int *out_dt) {
    if (op != IONP_ADD && op != IONP_MULTIPLY && op != IONP_MAXIMUM && op != IONP_MINIMUM)
        return set_error(IONP_ERR_VALUE, "reduce is only provided for add, multiply, maximum and minimum");
    int dt = dtype < 0 ? reduce_default_dtype(op, in_dt) : dtype;
    if (dt < 0 || dt >= IONP_NTYPES) return set_error(IONP_ERR_TYPE, "invalid dtype code");
    // the input is cast to the loop dtype unsafely (NumPy's reduce-family
    // default): a float input feeds an integer / bool loop element by
    // element with the C cast
    *out_dt = dt;
    return IONP_OK;
}

// reduce_axis's shape/dtype half; wst: the where= mask's strides broadcast
// to the input's shape (it joins the iterator's axis order), or null
static ionp_status plan_reduce_template(int32_t op, const ionp_array_desc *in, const int64_t *wst,
                                        const int32_t *axes, int32_t naxes, int32_t keepdims, int32_t dtype,
                                        ionp_array_desc *out_template) {
    ionp_status s;
    if ((s = check_desc(in, "array"))) return s;
    int odt;
    if ((s = reduce_dtype(op, in->dtype, dtype, &odt))) return s;
    bool mask[IONP_MAX_DIMS];
    if ((s = normalize_axes(axes, naxes, in->ndim, mask))) return s;
    // output layout: kept axes keep the input's iterator order (array.rs:910)
    int64_t st[2][IONP_MAX_DIMS];
    std::memcpy(st[0], in->strides, sizeof(int64_t) * in->ndim);
    if (wst) std::memcpy(st[1], wst, sizeof(int64_t) * in->ndim);
    int perm[IONP_MAX_DIMS];
    iter_axis_order(in->ndim, in->shape, wst ? 2 : 1, st, perm);
    int64_t kst[IONP_MAX_DIMS] = {0};
    int64_t acc = 1;
    for (int k = 0; k < in->ndim; k++) {
        const int ax = perm[k];
        if (mask[ax]) continue;
        kst[ax] = acc;
        acc *= in->shape[ax] > 0 ? in->shape[ax] : 1;
    }
    int64_t shape[IONP_MAX_DIMS], ost[IONP_MAX_DIMS];
    int nd = 0;
    for (int k = 0; k < in->ndim; k++) {
        if (mask[k]) {
            if (keepdims) {
                shape[nd] = 1;
                ost[nd++] = 0;
            }
        } else {
            shape[nd] = in->shape[k];
            ost[nd++] = kst[k];
        }
    }
    fill_template(out_template, odt, nd, shape, ost);
    return IONP_OK;
}

ionp_status ionp_plan_reduce(int32_t op, const ionp_array_desc *in, const int32_t *axes, int32_t naxes,
                             int32_t keepdims, int32_t dtype, ionp_array_desc *out_template) {
    return plan_reduce_template(op, in, nullptr, axes, naxes, keepdims, dtype, out_template);
}

// the where= operand: bool, broadcastable to the input's shape
// (broadcast_strides_to, shape.rs:113); its strides on the input's axes
static ionp_status where_strides(const ionp_array_desc *in, const ionp_array_desc *where, int64_t *wst) {
    ionp_status s;
    if ((s = check_desc(where, "where"))) return s;
    if (where->dtype != IONP_BOOL) return set_error(IONP_ERR_TYPE, "where= mask must have dtype bool");
    bool ok = where->ndim <= in->ndim;
    for (int i = 0; ok && i < where->ndim; i++) {
        const int64_t d = where->shape[i], t = in->shape[in->ndim - where->ndim + i];
        if (d != 1 && d != t) ok = false;
    }
    if (!ok)
        return set_error(IONP_ERR_BROADCAST, "operands could not be broadcast together with shapes " +
                                                 fmt_shape(where->shape, where->ndim) + " " +
                                                 fmt_shape(in->shape, in->ndim) + " ");
    broadcast_strides(where, in->ndim, in->shape, wst);
    return IONP_OK;
}

ionp_status ionp_plan_reduce_where(int32_t op, const ionp_array_desc *in, const ionp_array_desc *where,
                                   const int32_t *axes, int32_t naxes, int32_t keepdims, int32_t dtype,
                                   ionp_array_desc *out_template) {
    ionp_status s;
    if ((s = check_desc(in, "array"))) return s;
    int64_t wst[IONP_MAX_DIMS];
    if ((s = where_strides(in, where, wst))) return s;
    return plan_reduce_template(op, in, wst, axes, naxes, keepdims, dtype, out_template);
}

static ionp_status arg_normalize(const ionp_array_desc *in, int32_t axis, int *ax) {
    if (axis == IONP_AXIS_NONE) {
        *ax = -1;
        return IONP_OK;
    }
    const int eff = in->ndim > 1 ? in->ndim : 1;
    const int norm = axis < 0 ? axis + eff : axis;
    if (norm < 0 || norm >= eff)
        return set_error(IONP_ERR_AXIS, "axis " + std::to_string(axis) + " is out of bounds for array of dimension " +
                                            std::to_string(in->ndim));
    *ax = in->ndim == 0 ? -1 : norm;
    return IONP_OK;
}

ionp_status ionp_plan_argext(int32_t ismax, const ionp_array_desc *in, int32_t axis, int32_t keepdims,
                             ionp_array_desc *out_template) {
    ionp_status s;
    if ((s = check_desc(in, "array"))) return s;
    int ax;
    if ((s = arg_normalize(in, axis, &ax))) return s;
    int64_t shape[IONP_MAX_DIMS], ost[IONP_MAX_DIMS];
    int nd = 0;
    for (int k = 0; k < in->ndim; k++) {
        if (ax < 0 || k == ax) {
            if (keepdims) shape[nd++] = 1;
        } else {
            shape[nd++] = in->shape[k];
        }
    }
    int64_t acc = 1;
    for (int k = nd - 1; k >= 0; k--) {
        ost[k] = shape[k] == 1 ? 0 : acc;
        acc *= shape[k];
    }
    (void)ismax;
    fill_template(out_template, IONP_INT64, nd, shape, ost);
    return IONP_OK;
}

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------
ionp_status ionp_binary(int32_t op, const ionp_array_desc *a, const ionp_array_desc *b, const ionp_array_desc *out,
                        const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    ionp_config learned;
    AdviseTiming timing;
    if (!cfg) {
        const ionp_array_desc *ops[2] = {a, b};
        cfg = advise(IONP_CALL_BINARY, op, 2, ops, out, nullptr, 0, 0, as_stream(stream), &learned, &timing);
    }
    int loop, odt, nd;
    int64_t shape[IONP_MAX_DIMS];
    ionp_status s = binary_common(op, a, b, &loop, &odt, &nd, shape);
    if (s) return s;
    if ((s = check_desc(out, "out")) || (s = same_shape(out, nd, shape, op_name(op)))) return s;
    if (op >= IONP_LESS && out->dtype != IONP_BOOL)
        return set_error(IONP_ERR_TYPE, "comparison output must be bool");
    const ionp_array_desc *ins[2] = {a, b};
    EwArgs A;
    if ((s = plan_elementwise(op, 2, ins, out, loop, &A))) return s;
    if ((s = attach_fpe(A, op, loop, a->dtype, out->dtype))) return s;
    A.mixed_su = is_mixed_su(op, a->dtype, b->dtype);
    timing.begin();
    return timing.done(run_elementwise(A, cfg, as_stream(stream)));
}

ionp_status ionp_where(const ionp_array_desc *c, const ionp_array_desc *x, const ionp_array_desc *y,
                       const ionp_array_desc *out, const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    ionp_config learned;
    AdviseTiming timing;
    if (!cfg) {
        const ionp_array_desc *ops[3] = {c, x, y};
        cfg = advise(IONP_CALL_WHERE, 0, 3, ops, out, nullptr, 0, 0, as_stream(stream), &learned, &timing);
    }
    ionp_array_desc t;
    ionp_status s = ionp_plan_where(c, x, y, &t);
    if (s) return s;
    if ((s = check_desc(out, "out")) || (s = same_shape(out, t.ndim, t.shape, "where"))) return s;
    const ionp_array_desc *ins[3] = {c, x, y};
    EwArgs A;
    if ((s = plan_elementwise(OP_WHERE, 3, ins, out, out->dtype, &A))) return s;
    timing.begin();
    return timing.done(run_elementwise(A, cfg, as_stream(stream)));
}

ionp_status ionp_cast(const ionp_array_desc *in, const ionp_array_desc *out, const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    ionp_config learned;
    AdviseTiming timing;
    if (!cfg) {
        const ionp_array_desc *ops[1] = {in};
        cfg = advise(IONP_CALL_CAST, 0, 1, ops, out, nullptr, 0, 0, as_stream(stream), &learned, &timing);
    }
    ionp_status s;
    if ((s = check_desc(in, "array")) || (s = check_desc(out, "out"))) return s;
    if ((s = same_shape(out, in->ndim, in->shape, "astype"))) return s;
    const ionp_array_desc *ins[1] = {in};
    EwArgs A;
    if ((s = plan_elementwise(OP_CAST, 1, ins, out, out->dtype, &A))) return s;
    if ((s = attach_fpe(A, OP_CAST, out->dtype, in->dtype, out->dtype))) return s;
    timing.begin();
    return timing.done(run_elementwise(A, cfg, as_stream(stream)));
}

ionp_status ionp_reduce(int32_t op, const ionp_array_desc *in, const int32_t *axes, int32_t naxes, int32_t keepdims,
                        const void *initial, const ionp_array_desc *out, const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    if (fold_error_take()) return fold_failed();
    ionp_config learned;
    AdviseTiming timing;
    if (!cfg) {
        const ionp_array_desc *ops[1] = {in};
        cfg = advise(IONP_CALL_REDUCE, op, 1, ops, out, axes, naxes, keepdims, as_stream(stream), &learned, &timing);
    }
    ionp_array_desc t;
    ionp_status s;
    if ((s = check_desc(out, "out"))) return s;
    if ((s = ionp_plan_reduce(op, in, axes, naxes, keepdims, out->dtype, &t))) return s;
    if ((s = same_shape(out, t.ndim, t.shape, "reduce"))) return s;
    bool mask[IONP_MAX_DIMS];
    normalize_axes(axes, naxes, in->ndim, mask);
    ionp_array_desc c;
    if (needs_cast_copy(op, in->dtype, out->dtype)) {
        if ((s = cast_copy(in, out->dtype, &c, as_stream(stream)))) return s;
        in = &c;
    }
    ReducePlan p;
    if ((s = plan_reduce(op, in, mask, out, &p))) return s;
    timing.begin();
    return timing.done(run_reduce(p, g_bufsize, initial, cfg, as_stream(stream)));
}

ionp_status ionp_reduce_noalloc(int32_t op, const ionp_array_desc *in, const int32_t *axes, int32_t naxes,
                                int32_t keepdims, const void *initial, const ionp_array_desc *out,
                                const ionp_config *cfg, void *stream) {
    char *base;
    ionp_status s;
    if ((s = noalloc_base(&base))) return s;
    ScratchArena A;
    A.scratch = base;
    A.stage = base + kNoAllocPartials;
    A.slots = base + kNoAllocPartials + kNoAllocStage;
    A.size.scratch = kNoAllocPartials;
    A.size.stage = kNoAllocStage;
    A.size.slots = kNoAllocSlots;
    {
        // never the learned configuration: its first use on a device allocates
        static const ionp_config heuristic = {0, 0, 0, 0};
        ScratchScope scope(&A);
        s = ionp_reduce(op, in, axes, naxes, keepdims, initial, out, cfg ? cfg : &heuristic, stream);
    }
    if (A.short_) {
        auto mib = [](size_t b) { return std::to_string((b + (1u << 20) - 1) >> 20) + " MiB"; };
        std::string what;
        if (A.need.scratch > A.size.scratch) what = "partials " + mib(A.need.scratch) + " (maximum " + mib(A.size.scratch) + ")";
        else if (A.need.stage > A.size.stage) what = "a float16 input converted to " + mib(A.need.stage) + " (maximum " + mib(A.size.stage) + ")";
        else what = "fold slots " + mib(A.need.slots) + " (maximum " + mib(A.size.slots) + ")";
        return set_error(IONP_ERR_VALUE, "ionp_reduce_noalloc: this reduction needs " + what +
                                             " of its fixed scratch; use ionp_reduce");
    }
    return s;
}

ionp_status ionp_reduce_where(int32_t op, const ionp_array_desc *in, const ionp_array_desc *where,
                              const int32_t *axes, int32_t naxes, int32_t keepdims, const void *initial,
                              const ionp_array_desc *out, const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    ionp_array_desc t;
    ionp_status s;
    if ((s = check_desc(out, "out"))) return s;
    if ((s = ionp_plan_reduce_where(op, in, where, axes, naxes, keepdims, out->dtype, &t))) return s;
    if ((s = same_shape(out, t.ndim, t.shape, "reduce"))) return s;
    // reduce_axis_masked (ufunc.rs:7172): no identity and no initial= is an
    // error whatever the sizes
    if ((op == IONP_MAXIMUM || op == IONP_MINIMUM) && !initial)
        return set_error(IONP_ERR_VALUE, std::string("reduction operation '") + op_name(op) +
                                             "' does not have an identity, so to use a where mask one has to "
                                             "specify 'initial'");
    int64_t wst[IONP_MAX_DIMS];
    if ((s = where_strides(in, where, wst))) return s;
    bool mask[IONP_MAX_DIMS];
    normalize_axes(axes, naxes, in->ndim, mask);
    ionp_array_desc c;
    if (needs_cast_copy(op, in->dtype, out->dtype)) {  // a K-order copy keeps the iterator's axis order
        if ((s = cast_copy(in, out->dtype, &c, as_stream(stream)))) return s;
        in = &c;
    }
    ReducePlan p;
    if ((s = plan_reduce(op, in, mask, out, &p, wst, (const uint8_t *)elem0(where)))) return s;
    return run_reduce(p, g_bufsize, initial, cfg, as_stream(stream));
}

static ionp_status acc_axis(const ionp_array_desc *in, int32_t axis, int *ax) {
    if (in->ndim == 0) return set_error(IONP_ERR_VALUE, "cannot accumulate on a scalar");
    const int norm = axis < 0 ? axis + in->ndim : axis;
    if (norm < 0 || norm >= in->ndim)
        return set_error(IONP_ERR_AXIS, "axis " + std::to_string(axis) + " is out of bounds for array of dimension " +
                                            std::to_string(in->ndim));
    *ax = norm;
    return IONP_OK;
}

ionp_status ionp_plan_accumulate(int32_t op, const ionp_array_desc *in, int32_t axis, int32_t dtype,
                                 ionp_array_desc *out_template) {
    ionp_status s;
    if ((s = check_desc(in, "array"))) return s;
    int odt, ax;
    if ((s = reduce_dtype(op, in->dtype, dtype, &odt))) return s;
    if ((s = acc_axis(in, axis, &ax))) return s;
    // the input's K order (the ordinary one-operand layout rule, as for casts)
    int64_t st[1][IONP_MAX_DIMS], ost[IONP_MAX_DIMS];
    std::memcpy(st[0], in->strides, sizeof(int64_t) * in->ndim);
    korder_strides(1, st, in->ndim, in->shape, ost);
    fill_template(out_template, odt, in->ndim, in->shape, ost);
    return IONP_OK;
}

ionp_status ionp_accumulate(int32_t op, const ionp_array_desc *in, int32_t axis, const ionp_array_desc *out,
                            const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    ionp_status s;
    ionp_array_desc t;
    if ((s = check_desc(out, "out"))) return s;
    if ((s = ionp_plan_accumulate(op, in, axis, out->dtype, &t))) return s;
    if ((s = same_shape(out, t.ndim, t.shape, "accumulate"))) return s;
    int ax = 0;
    acc_axis(in, axis, &ax);
    hipStream_t st = as_stream(stream);
    if (in->dtype == out->dtype) return run_accumulate(op, in, ax, out, cfg, st);
    // another input dtype: cast first (dense, K order), then scan
    ionp_array_desc c;
    if ((s = cast_copy(in, out->dtype, &c, st))) return s;
    return run_accumulate(op, &c, ax, out, cfg, st);
}

// reduceat's shape/dtype half (ionp-py lib.rs:10127 -> reduceat_binary_axis,
// ufunc.rs:8460): axis normalised, output dtype by the reduce rule, layout
// in NumPy's K order over (input without `axis`, indices along `axis`)
static ionp_status rat_plan(int32_t op, const ionp_array_desc *in, int32_t axis, int64_t nidx, int32_t dtype,
                            ionp_array_desc *out_template, int *ax_out) {
    ionp_status s;
    if ((s = check_desc(in, "array"))) return s;
    if (in->ndim == 0) return set_error(IONP_ERR_TYPE, "cannot reduceat on a scalar");
    int ax;
    if ((s = acc_axis(in, axis, &ax))) return s;
    if (nidx < 0) return set_error(IONP_ERR_VALUE, "negative number of reduceat indices");
    int odt;
    if ((s = reduce_dtype(op, in->dtype, dtype, &odt))) return s;
    // a float loop reads the input as is (a cast would go through NumPy's
    // buffers and change the pairwise tree)
    if (is_float_dt(odt) && odt != in->dtype)
        return set_error(IONP_ERR_UNSUPPORTED, "float reduceat with a dtype other than the input's");
    int64_t shape[IONP_MAX_DIMS], st[2][IONP_MAX_DIMS], ost[IONP_MAX_DIMS];
    for (int k = 0; k < in->ndim; k++) {
        shape[k] = k == ax ? nidx : in->shape[k];
        st[0][k] = k == ax ? 0 : in->strides[k];
        st[1][k] = k == ax ? 1 : 0;
    }
    korder_strides(2, st, in->ndim, shape, ost);
    fill_template(out_template, odt, in->ndim, shape, ost);
    *ax_out = ax;
    return IONP_OK;
}

ionp_status ionp_plan_reduceat(int32_t op, const ionp_array_desc *in, int32_t axis, int64_t nidx, int32_t dtype,
                               ionp_array_desc *out_template) {
    int ax;
    return rat_plan(op, in, axis, nidx, dtype, out_template, &ax);
}

ionp_status ionp_reduceat(int32_t op, const ionp_array_desc *in, const int64_t *indices, int64_t nidx, int32_t axis,
                          const ionp_array_desc *out, const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    ionp_status s;
    ionp_array_desc t;
    if ((s = check_desc(out, "out"))) return s;
    if (in->ndim == 0) return set_error(IONP_ERR_TYPE, "cannot reduceat on a scalar");
    int ax;
    if ((s = acc_axis(in, axis, &ax))) return s;
    // indices in call order, before any loop question (validate_reduceat_indices, ufunc.rs:8426)
    const int64_t n = in->shape[ax];
    if (nidx > 0 && !indices) return set_error(IONP_ERR_VALUE, "reduceat indices missing");
    for (int64_t i = 0; i < nidx; i++)
        if (indices[i] < 0 || indices[i] >= n)
            return set_error(IONP_ERR_INDEX, "index " + std::to_string(indices[i]) + " out-of-bounds in " +
                                                 op_name(op) + ".reduceat [0, " + std::to_string(n) + ")");
    if ((s = rat_plan(op, in, axis, nidx, out->dtype, &t, &ax))) return s;
    if ((s = same_shape(out, t.ndim, t.shape, "reduceat"))) return s;
    RatArgs A;
    std::memset(&A, 0, sizeof A);
    A.op = op;
    A.in_dt = in->dtype;
    A.out_dt = out->dtype;
    A.len = n;
    A.in_st = in->strides[ax];
    A.out_st = out->strides[ax];
    A.nidx = nidx;
    A.nfib = 1;
    // fiber dims innermost first by the input's strides
    int64_t st[1][IONP_MAX_DIMS];
    std::memcpy(st[0], in->strides, sizeof(int64_t) * in->ndim);
    int perm[IONP_MAX_DIMS];
    iter_axis_order(in->ndim, in->shape, 1, st, perm);
    for (int k = 0; k < in->ndim; k++) {
        const int d = perm[k];
        if (d == ax || in->shape[d] == 1) continue;
        A.fshape[A.nf] = in->shape[d];
        A.fin[A.nf] = in->strides[d];
        A.fout[A.nf] = out->strides[d];
        A.nf++;
        A.nfib *= in->shape[d];
    }
    A.in = elem0(in);
    A.out = (char *)elem0(out);
    return run_reduceat(A, indices, cfg, as_stream(stream));
}

ionp_status ionp_argext(int32_t ismax, const ionp_array_desc *in, int32_t axis, int32_t keepdims,
                        const ionp_array_desc *out, const ionp_config *cfg, void *stream) {
    if (ionp_status cs = check_cfg(cfg)) return cs;
    ionp_array_desc t;
    ionp_status s;
    if ((s = check_desc(out, "out"))) return s;
    if ((s = ionp_plan_argext(ismax, in, axis, keepdims, &t))) return s;
    if ((s = same_shape(out, t.ndim, t.shape, "argext"))) return s;
    if (out->dtype != IONP_INT64) return set_error(IONP_ERR_TYPE, "argmax/argmin output must be int64");
    int ax;
    arg_normalize(in, axis, &ax);
    ionp_array_desc c;
    if (in->dtype == IONP_FLOAT16) {  // float16 -> float32 is exact and keeps the order, NaNs included
        if ((s = cast_copy(in, IONP_FLOAT32, &c, as_stream(stream)))) return s;
        in = &c;
    }
    ArgArgs A;
    std::memset(&A, 0, sizeof A);
    A.in_dt = in->dtype;
    A.ismax = ismax ? 1 : 0;
    A.in = elem0(in);
    A.out = (char *)elem0(out);
    const int64_t total = desc_size(in);
    if (ax < 0) {
        A.len = total;
        A.nout = 1;
        if (total == 0)
            return set_error(IONP_ERR_VALUE, std::string("attempt to get ") + (ismax ? "argmax" : "argmin") +
                                                 " of an empty sequence");
        // C-order ravel, innermost first, coalesced where the strides allow
        for (int k = in->ndim - 1; k >= 0; k--) {
            if (in->shape[k] == 1) continue;
            if (A.nl > 0 && in->strides[k] == A.l_st[A.nl - 1] * A.l_shape[A.nl - 1]) {
                A.l_shape[A.nl - 1] *= in->shape[k];
            } else {
                A.l_shape[A.nl] = in->shape[k];
                A.l_st[A.nl++] = in->strides[k];
            }
        }
        if (A.nl == 0) {
            A.nl = 1;
            A.l_shape[0] = 1;
            A.l_st[0] = 0;
        }
    } else {
        A.len = in->shape[ax];
        A.nl = 1;
        A.l_shape[0] = A.len;
        A.l_st[0] = in->strides[ax];
        A.nout = 1;
        // kept axes innermost first (reverse C order) with output strides
        int j = out->ndim - 1;
        for (int k = in->ndim - 1; k >= 0; k--) {
            if (k == ax) {
                if (keepdims) j--;
                continue;
            }
            const int64_t n = in->shape[k];
            const int64_t ost = out->strides[j--];
            A.nout *= n;
            if (n == 1) continue;
            A.k_shape[A.nk] = n;
            A.k_in_st[A.nk] = in->strides[k];
            A.k_out_st[A.nk++] = ost;
        }
        if (A.len == 0 && A.nout > 0)
            return set_error(IONP_ERR_VALUE, std::string("attempt to get ") + (ismax ? "argmax" : "argmin") +
                                                 " of an empty sequence");
    }
    return run_argext(A, cfg, as_stream(stream));
}

}  // extern "C"
