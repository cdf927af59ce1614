/*
 * ionp_hip.h -- C ABI of the MI355X (gfx950) ufunc / reduction backend.
 *
 * Plain C: fixed-size descriptors, device pointers, status codes.  No HIP,
 * torch or C++ types cross this boundary; `stream` is a hipStream_t passed
 * as void* (NULL = the default stream).  A Rust `ionp-hip` crate binds it
 * with `extern "C"` verbatim; the Python package anionpy_amd binds it with
 * ctypes.
 *
 * Each entry point names the reference function it replaces.  Paths are in
 * the anionpy tree (Son0fSun/anionpy).
 *
 * Conventions
 *   - Strides and offsets are in ELEMENTS (ionp-core/src/shape.rs:1-5,
 *     ionp-core/src/array.rs:24-29); negative and zero strides are allowed.
 *   - Callers allocate every output.  The plan_* calls fill an output
 *     template (dtype, shape, NumPy K-order strides) without touching the
 *     device, so the caller can allocate it.
 *   - Results are bit-identical to NumPy 2.2.6 on x86-64 (the reference's
 *     own oracle), including add.reduce's pairwise tree per 8192-element
 *     iterator buffer window.
 *   - Calls are stream-ordered and thread-safe; the last error message is
 *     thread-local.
 */
#ifndef IONP_HIP_H
#define IONP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IONP_MAX_DIMS 16
#define IONP_AXIS_NONE (-1000000)

/* dtype codes (ionp-core/src/dtype.rs DType, numeric subset) */
typedef enum {
    IONP_BOOL = 0,
    IONP_INT8 = 1,
    IONP_UINT8 = 2,
    IONP_INT16 = 3,
    IONP_UINT16 = 4,
    IONP_INT32 = 5,
    IONP_UINT32 = 6,
    IONP_INT64 = 7,
    IONP_UINT64 = 8,
    IONP_FLOAT32 = 9,
    IONP_FLOAT64 = 10,
    IONP_FLOAT16 = 11, /* IEEE binary16 (NumPy float16) */
    IONP_NTYPES = 12
} ionp_dtype;

/* binary ufuncs (ionp-core/src/ufunc.rs:43-65 BinaryOp, subset) */
typedef enum {
    IONP_ADD = 0,
    IONP_SUBTRACT = 1,
    IONP_MULTIPLY = 2,
    IONP_TRUE_DIVIDE = 3,
    IONP_MAXIMUM = 4,
    IONP_MINIMUM = 5,
    IONP_LESS = 6,
    IONP_LESS_EQUAL = 7,
    IONP_GREATER = 8,
    IONP_GREATER_EQUAL = 9,
    IONP_EQUAL = 10,
    IONP_NOT_EQUAL = 11,
    IONP_NOPS = 12
} ionp_op;

/* status codes; 1..7 map 1:1 onto IonpError (ionp-core/src/error.rs:8-42) */
typedef enum {
    IONP_OK = 0,
    IONP_ERR_BROADCAST = 1, /* ValueError: operands could not be broadcast ... */
    IONP_ERR_RESHAPE = 2,   /* ValueError */
    IONP_ERR_INDEX = 3,     /* IndexError */
    IONP_ERR_TYPE = 4,      /* TypeError */
    IONP_ERR_VALUE = 5,     /* ValueError */
    IONP_ERR_NO_LOOP = 6,   /* UFuncTypeError (TypeError subclass) */
    IONP_ERR_AXIS = 7,      /* AxisError (ValueError and IndexError) */
    IONP_ERR_DEVICE = 20,   /* HIP runtime failure */
    IONP_ERR_UNSUPPORTED = 21
} ionp_status;

/* Strided array view in device memory (NdArray, ionp-core/src/array.rs:24-29). */
typedef struct {
    int32_t dtype; /* ionp_dtype */
    int32_t ndim;
    int64_t shape[IONP_MAX_DIMS];
    int64_t strides[IONP_MAX_DIMS]; /* elements */
    int64_t offset;                 /* elements from `data` to element (0,...,0) */
    void *data;                     /* device pointer */
} ionp_array_desc;

/* Launch configuration.  Zero fields mean "use the built-in heuristic".
 * Every configuration computes bit-identical results: it changes only which
 * thread handles which element, never the arithmetic order. */
typedef struct {
    int32_t block;         /* threads per workgroup: 64, 128, 192 or 256 */
    int32_t vec;           /* elements per vector access: 1, 2, 4, 8 */
    int32_t blocks_per_cu; /* grid size = blocks_per_cu * CUs, at most the work items (grid-stride) */
    int32_t split;         /* reduce: outputs per thread (narrow) / unroll */
} ionp_config;

/* ---- library state --------------------------------------------------- */

/* Initialise on the current device (idempotent).  Optional: every entry
 * point initialises lazily. */
ionp_status ionp_init(void);

/* Text of the last error on this thread, NumPy-verbatim where NumPy has a
 * message (IonpError's Display, ionp-core/src/error.rs:64-90). */
const char *ionp_last_error(void);

/* ABI version (major << 16 | minor). */
int32_t ionp_version(void);

/* NumPy's iterator buffer size used to window add.reduce (np.getbufsize(),
 * default 8192).  Values 16..8192 accepted. */
ionp_status ionp_set_bufsize(int64_t elements);
int64_t ionp_get_bufsize(void);

/* ---- type resolution (host only) ------------------------------------- */

/* promote_dtype (ionp-core/src/dtype.rs:297). */
ionp_status ionp_result_type(int32_t a, int32_t b, int32_t *out);

/* binary_out_dtype (ionp-core/src/ufunc.rs:1101): loop and output dtype. */
ionp_status ionp_binary_types(int32_t op, int32_t a, int32_t b, int32_t *loop, int32_t *out);

/* casting rules (NumPy's casting= keyword) */
typedef enum {
    IONP_CAST_NO = 0,
    IONP_CAST_EQUIV = 1,
    IONP_CAST_SAFE = 2,
    IONP_CAST_SAME_KIND = 3,
    IONP_CAST_UNSAFE = 4
} ionp_casting;

/* dtype::can_cast (ionp-core/src/dtype.rs): *ok = 1 when `from` casts to
 * `to` under `casting`. */
ionp_status ionp_can_cast(int32_t from, int32_t to, int32_t casting, int32_t *ok);

/* Ufunc.__call__'s dtype= resolution (ionp-py/src/lib.rs:9568-9804 with
 * resolve_binary_output_dtype, ionp-core/src/ufunc.rs:1031): the loop dtype
 * the inputs are cast to before the call.  Errors as NumPy's: TypeError
 * "No loop matching the specified signature and casting was found for
 * ufunc <name>" (IONP_ERR_TYPE), UFuncTypeError "Cannot cast ufunc '<name>'
 * input <i> from dtype('<a>') to dtype('<b>') with casting rule '<rule>'"
 * (IONP_ERR_NO_LOOP). */
ionp_status ionp_resolve_dtype(int32_t op, int32_t a, int32_t b, int32_t requested, int32_t casting,
                               int32_t *loop);

/* ---- planning (host only: fills an output template) ------------------ */

/* binary_op's shape/dtype/layout half (ufunc.rs:3628-3695, shape.rs:54,
 * lib.rs:1290 apply_ufunc_order 'K'). */
ionp_status ionp_plan_binary(int32_t op, const ionp_array_desc *a, const ionp_array_desc *b,
                             ionp_array_desc *out_template);

/* where_select's shape/dtype/layout half (sort.rs:1869-1914). */
ionp_status ionp_plan_where(const ionp_array_desc *cond, const ionp_array_desc *x,
                            const ionp_array_desc *y, ionp_array_desc *out_template);

/* NdArray::cast_to's layout half (array.rs:393-429), order 'K'. */
ionp_status ionp_plan_cast(const ionp_array_desc *in, int32_t dtype, ionp_array_desc *out_template);

/* reduce_axis's shape/dtype half (ufunc.rs:7551-7617 normalize_reduce_axes,
 * 7634-7825, array.rs:910 reduction output layout).  naxes < 0 means
 * axis=None; dtype < 0 means the default reduce dtype. */
ionp_status ionp_plan_reduce(int32_t op, const ionp_array_desc *in, const int32_t *axes, int32_t naxes,
                             int32_t keepdims, int32_t dtype, ionp_array_desc *out_template);

/* reduce_axis's shape/dtype half with a where= mask (ufunc.rs:7634 with
 * `where_mask`): `where` is a bool array broadcastable to the input's
 * shape (broadcast_strides_to, shape.rs:113); it takes part in the
 * iterator's axis order, so it can change the output's layout. */
ionp_status ionp_plan_reduce_where(int32_t op, const ionp_array_desc *in, const ionp_array_desc *where,
                                   const int32_t *axes, int32_t naxes, int32_t keepdims, int32_t dtype,
                                   ionp_array_desc *out_template);

/* argext's shape half (ufunc.rs:8131-8206); axis = IONP_AXIS_NONE for the
 * flattened form. */
ionp_status ionp_plan_argext(int32_t ismax, const ionp_array_desc *in, int32_t axis, int32_t keepdims,
                             ionp_array_desc *out_template);

/* ---- execution (device) ---------------------------------------------- */

/* Replaces ionp_core::ufunc::binary_op (ufunc.rs:3628).  `out` must have
 * the broadcast shape and the binary_out_dtype; any strides. */
ionp_status ionp_binary(int32_t op, const ionp_array_desc *a, const ionp_array_desc *b,
                        const ionp_array_desc *out, const ionp_config *cfg, void *stream);

/* Replaces ionp_core::sort::where_select (sort.rs:1869). */
ionp_status ionp_where(const ionp_array_desc *cond, const ionp_array_desc *x, const ionp_array_desc *y,
                       const ionp_array_desc *out, const ionp_config *cfg, void *stream);

/* Replaces NdArray::cast_to (array.rs:393) / Buffer::cast_to (buffer.rs:210). */
ionp_status ionp_cast(const ionp_array_desc *in, const ionp_array_desc *out, const ionp_config *cfg,
                      void *stream);

/* Replaces ionp_core::ufunc::reduce_axis (ufunc.rs:7634) for op in
 * {ADD, MULTIPLY, MAXIMUM, MINIMUM}: `out` as given by ionp_plan_reduce (its
 * dtype selects the accumulation dtype).  Float multiply is one running
 * product per output in iterator order (no pairwise tree).  `initial`:
 * NULL, or one element of the output dtype in host memory (reduce_axis's
 * `initial`): the fold starts from it (add: acc = initial; acc +=
 * pairwise(window) per window; multiply: acc = initial), maximum and
 * minimum feed every element to the loop instead of seeding with the
 * first, and an empty reduction returns it (else add's 0, multiply's 1). */
ionp_status ionp_reduce(int32_t op, const ionp_array_desc *in, const int32_t *axes, int32_t naxes,
                        int32_t keepdims, const void *initial, const ionp_array_desc *out, const ionp_config *cfg,
                        void *stream);

/* No counterpart in the reference: ionp_reduce for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, results and
 * errors as ionp_reduce, but every scratch buffer (reduction partials, a
 * float16 input converted for a float32 / float64 loop, the in-launch
 * folds' slots) comes from one fixed 8 MiB array per GPU that the library
 * holds until the process ends (its address is looked up by ionp_init, or
 * by the first call on a device), so no call allocates or frees anything.
 * Limits: 4 MiB of partials (2^19 float64 window partials), 2 MiB of
 * converted input (2^19 float16 elements into float32), 2 MiB of fold
 * slots (2^18 float64 windows: one-output sums of up to 2^31 elements); a
 * reduction needing more is rejected with IONP_ERR_VALUE before its output
 * is written.  All calls on
 * one GPU share the array: calls that overlap in time on one GPU (other
 * streams or threads) race. */
ionp_status ionp_reduce_noalloc(int32_t op, const ionp_array_desc *in, const int32_t *axes, int32_t naxes,
                                int32_t keepdims, const void *initial, const ionp_array_desc *out,
                                const ionp_config *cfg, void *stream);

/* Replaces reduce_axis (ufunc.rs:7634) when `where_mask` is given
 * (reduce_axis_masked, ufunc.rs:7172), as NumPy 2.2.6 evaluates it: the
 * fold starts from `initial` (else add's 0, multiply's 1; maximum and
 * minimum without `initial` are an error whatever the sizes), and inside
 * every iterator buffer window the reduce loop runs once per maximal run
 * of True mask values (float add: acc = acc + pairwise(run)).  `out` as
 * given by ionp_plan_reduce_where. */
ionp_status ionp_reduce_where(int32_t op, const ionp_array_desc *in, const ionp_array_desc *where,
                              const int32_t *axes, int32_t naxes, int32_t keepdims, const void *initial,
                              const ionp_array_desc *out, const ionp_config *cfg, void *stream);

/* Replaces ionp_core::ufunc::argext (ufunc.rs:8131): int64 indices. */
ionp_status ionp_argext(int32_t ismax, const ionp_array_desc *in, int32_t axis, int32_t keepdims,
                        const ionp_array_desc *out, const ionp_config *cfg, void *stream);

/* Replaces ionp_core::ufunc::accumulate_axis (ufunc.rs:7967) for op in
 * {ADD, MULTIPLY, MAXIMUM, MINIMUM}: along `axis` (negative counts from the
 * end), out[0] = x[0], out[i] = op(out[i-1], x[i]) in the output dtype (the
 * reduce dtype rule: small integers and bool add and multiply in int64 /
 * uint64; or an explicit dtype).  Float add and multiply are the sequential
 * running sum / product.  The output
 * template has the input's shape in its K order. */
ionp_status ionp_plan_accumulate(int32_t op, const ionp_array_desc *in, int32_t axis, int32_t dtype,
                                 ionp_array_desc *out_template);
ionp_status ionp_accumulate(int32_t op, const ionp_array_desc *in, int32_t axis, const ionp_array_desc *out,
                            const ionp_config *cfg, void *stream);

/* Replaces reduceat_binary_axis (ufunc.rs:8460; Python entry
 * ionp-py/src/lib.rs:10127): along `axis`, output k of every fiber is
 * x[indices[k]] when indices[k+1] <= indices[k] (the last index runs to
 * the axis end), else the reduce loop seeded with x[indices[k]] over the
 * rest of the segment as one call, as NumPy 2.2.6 evaluates it (float
 * add: x[i] + pairwise(rest), no buffer windows; multiply sequential;
 * max/min the x86 loop model).  `indices`: nidx host int64 values, each in
 * [0, n) (else IONP_ERR_INDEX with NumPy's text).  The output template
 * has the input's shape with `axis` of length nidx, in NumPy's K order.
 * A float output dtype other than the input's is not provided (21). */
ionp_status ionp_plan_reduceat(int32_t op, const ionp_array_desc *in, int32_t axis, int64_t nidx, int32_t dtype,
                               ionp_array_desc *out_template);
ionp_status ionp_reduceat(int32_t op, const ionp_array_desc *in, const int64_t *indices, int64_t nidx, int32_t axis,
                          const ionp_array_desc *out, const ionp_config *cfg, void *stream);

/* ---- IEEE exception flags (np.seterr / np.errstate) ------------------- */

/* The fpe_check_after_call side channel (ionp-py/src/lib.rs:9077-9196,
 * ionp-core/src/fpe.rs:173-360), computed on the device instead of by a
 * host rescan.  Tracking is per calling thread, as the reference keeps its
 * FPE state per thread (ionp-py/src/fpstate.rs:103): while it is on for a
 * thread, that thread's ionp_binary and ionp_cast calls OR the IEEE flags
 * their elements raise -- as NumPy 2.2.6 reads them from the x86 status
 * register after a loop -- into the thread's own flag word on the current
 * GPU.  Bits: 1 divide by zero, 2 overflow, 4 underflow (tiny after
 * rounding and inexact), 8 invalid.  Float add / subtract / multiply /
 * true_divide and casts from a float dtype raise flags; comparisons,
 * maximum / minimum, where and integer loops never do.  Reductions,
 * batches (ionp_batch_run) and ionp_autotune are never tracked.
 * Switching tracking on or off waits for the device and clears the
 * thread's word, so no flag outlives the state that raised it. */
ionp_status ionp_fpe_track(int32_t on);
/* Waits for `stream`, returns the calling thread's flags on the current
 * GPU and clears them. */
ionp_status ionp_fpe_fetch(void *stream, uint32_t *flags);

/* ---- device failures reported without a synchronisation -------------- */

/* Waits for `stream`, then reports a failure a launch recorded on the
 * device: the in-launch fold of a one-output add.reduce (over 64K
 * elements) waits a bounded time for window partials, and if a producer
 * workgroup is never scheduled (its CUs held by another process) it gives
 * up, writes NaN as that output and records it here (IONP_ERR_DEVICE).
 * ionp_reduce also returns such a recorded failure at its next call. */
ionp_status ionp_check(void *stream);

/* Test hook (no counterpart in the reference): the folds of later
 * add.reduce calls treat window `window` (-1: none) as never filled and
 * give up after `polls` polls of one unfilled slot (default 2^26), so the
 * failure path above runs while every kernel still exits normally. */
ionp_status ionp_debug_fold_stall(int64_t window, uint32_t polls);

/* ---- batching: many small calls, one launch --------------------------- */
/* No counterpart in the reference (it runs one call at a time). */

typedef enum {
    IONP_CALL_BINARY = 0,
    IONP_CALL_WHERE = 1,
    IONP_CALL_CAST = 2,
    IONP_CALL_REDUCE = 3
} ionp_call_kind;

typedef struct {
    int32_t kind;     /* ionp_call_kind */
    int32_t op;       /* ionp_op for BINARY / REDUCE */
    int32_t naxes;    /* REDUCE: < 0 means axis=None */
    int32_t keepdims; /* REDUCE */
    int32_t axes[IONP_MAX_DIMS];
    ionp_array_desc in[3]; /* BINARY: a,b; WHERE: cond,x,y; CAST/REDUCE: in */
    ionp_array_desc out;
} ionp_call;

typedef struct ionp_batch ionp_batch;

typedef enum {
    IONP_BATCH_PERSISTENT = 0, /* one persistent-kernel launch over a tile queue */
    IONP_BATCH_GRAPH = 1,      /* per-call launches captured once into a hipGraph */
    IONP_BATCH_PER_CALL = 2    /* one stream launch sequence per call */
} ionp_batch_mode;

/* Tile sizes of a batch's persistent kernel (zero fields: the defaults).
 * Like ionp_config they move work between workgroups only: results are
 * bit-identical whatever is chosen. */
typedef struct {
    int64_t ew_tile; /* elements per elementwise tile: a multiple of 1024 (default 16384) */
    int64_t int_seg; /* elements per integer-reduce segment (default 8192) */
    int32_t claim;   /* consecutive tiles a workgroup claims at once (1..64, default 4) */
    int32_t reserved;
} ionp_batch_config;

/* Validate and plan `n` calls, pack their descriptors and tile table into
 * device memory.  The calls' device buffers must outlive the batch.  Calls
 * are independent: no call's output may overlap another call's operands.
 * In PERSISTENT mode float16 calls run as their own single-call launches,
 * in stream order ahead of the persistent kernel. */
ionp_status ionp_batch_create(const ionp_call *calls, int32_t n, void *stream, ionp_batch **out);
/* ionp_batch_create with tile sizes (null: the defaults). */
ionp_status ionp_batch_create_ex(const ionp_call *calls, int32_t n, const ionp_batch_config *config, void *stream,
                                 ionp_batch **out);
ionp_status ionp_batch_run(ionp_batch *batch, int32_t mode, void *stream);
/* number of tiles / bytes moved (algorithmic) by one run */
ionp_status ionp_batch_info(const ionp_batch *batch, int64_t *ntiles, int64_t *bytes);
void ionp_batch_destroy(ionp_batch *batch);

/* Times the batch's PERSISTENT run under each candidate tile schedule:
 * times_us[j] is the fastest of `reps` runs with cands[j] (a schedule the
 * batch cannot use: a negative time).  The calls' outputs are overwritten;
 * each candidate builds and frees its own batch. */
ionp_status ionp_batch_autotune(const ionp_call *calls, int32_t n, const ionp_batch_config *cands, int32_t ncand,
                                int32_t reps, float *times_us, void *stream);

/* ---- autotuning (device) ----------------------------------------------- */
/* No counterpart in the reference (its CPU loops have no launch knobs). */

/* Times every call under every configuration with HIP events on `stream`:
 * times_us[i * ncfg + j] is the fastest of `reps` runs of call i with
 * cfgs[j] (a configuration a call cannot use is reported as a negative
 * time).  The calls' outputs are overwritten. */
ionp_status ionp_autotune(const ionp_call *calls, int32_t n, const ionp_config *cfgs, int32_t ncfg, int32_t reps,
                          float *times_us, void *stream);

/* ---- learned launch configuration (the advisor) ----------------------- */
/* No counterpart in the reference.
 *
 * With a model installed, ionp_binary / ionp_where / ionp_cast /
 * ionp_reduce called with cfg == NULL run the configuration the cost model
 * proposes for the call's layout (a zero ionp_config still means the
 * built-in heuristic).  Proposals are cached per layout (its feature vector).
 * A layout's first call runs the heuristic and queues the layout; the MLP's
 * forward kernels evaluate the queue on that call's stream (at most 256
 * layouts per pass, one pass in flight per device) and a later call
 * installs the result once the pass has finished; a call on a stream under
 * graph capture only reads the cache (no event query, pass or timing enters
 * the capture),
 * so no call waits for the model.  Each proposal is then verified on live
 * calls: three calls of the heuristic (the layout's early calls, not its
 * first) and three of the proposal are timed with HIP events, read back
 * without waiting by later calls, and the layout keeps whichever ran
 * faster.  Results are
 * bit-identical whatever runs.  ionp_reduce_noalloc, batches and
 * ionp_autotune always run the heuristic.  Thread-safe, except that
 * ionp_advisor_set must not run while other threads make calls; the cache
 * is shared by all devices. */

/* Installs a model (nlayers > 0) or removes it (nlayers == 0: every call
 * runs the heuristic).  sizes[0..nlayers]: layer widths, sizes[0] ==
 * IONP_NFEATURES; params: per layer W (sizes[l] x sizes[l+1], row-major)
 * then b; heads[nheads]: values per knob (block, vec, blocks_per_cu,
 * split), and with nheads == 5 a fallback head of two values whose second
 * (value 1) means "run the built-in heuristic"; values: the knob values,
 * head after head (sum(heads) == sizes[nlayers]).
 * Clears the cache; device copies are made at each device's first pass. */
ionp_status ionp_advisor_set(const int32_t *sizes, int32_t nlayers, const float *params, const int32_t *heads,
                             int32_t nheads, const int32_t *values);
/* Evaluates the model for these calls' layouts now (batched passes on
 * `stream`, waiting for them), so their first calls already run the
 * learned configuration. */
ionp_status ionp_advisor_prime(const ionp_call *calls, int32_t n, void *stream);
/* stats[0]: calls that ran the advisor's configuration for their layout
 * (the proposal, or the heuristic a trial kept), [1]: calls that ran the
 * heuristic while their layout's proposal was pending or under trial, [2]:
 * layouts evaluated, [3]: layouts queued or in flight, [4]: trials that
 * kept the proposal, [5]: trials that kept the heuristic, [6]: trials in
 * progress (stats holds 7 entries). */
ionp_status ionp_advisor_stats(int64_t *stats);
/* The cached proposal for a call's layout (*found = 0 if none yet). */
ionp_status ionp_advisor_lookup(const ionp_call *call, ionp_config *cfg, int32_t *found);

/* ---- layout features for the cost model (host only) ------------------ */
/* No counterpart in the reference. */

#define IONP_NFEATURES 32

/* Feature vector of a call's coalesced layout: rank, log2 dims and total
 * size, per-operand contiguity / broadcast / negative-stride bits along the
 * innermost coalesced axis, the output's contiguity there, input dtype and
 * op class one-hots, input and output item sizes, log2 of the largest inner
 * stride, log2 of the call's bytes.  Entries 30-31 are zero (reserved). */
ionp_status ionp_call_features(const ionp_call *call, float *features);

/* ---- MLP cost model kernels (device) ---------------------------------- */
/* No counterpart in the reference.  fp32, row-major, hidden width 64. */

/* y[n, out] = relu?(x[n, in] @ w[in, out] + b[out]) */
ionp_status ionp_mlp_linear_fwd(const float *x, const float *w, const float *b, float *y, int32_t n,
                                int32_t in, int32_t out, int32_t relu, void *stream);
/* Given dy (already masked by the ReLU derivative if relu was used), y
 * (forward output, for the ReLU mask) -> dx, dw, db. dx may be NULL. */
ionp_status ionp_mlp_linear_bwd(const float *x, const float *w, const float *y, const float *dy,
                                float *dx, float *dw, float *db, int32_t n, int32_t in, int32_t out,
                                int32_t relu, void *stream);
/* Softmax cross-entropy over `heads` groups of logits: logits[n, total],
 * head h occupies [off[h], off[h+1]) (off and labels[n, heads] in device
 * memory); writes dlogits (gradient of the mean loss) and loss, a device
 * buffer of n + 1 floats: loss[0] = mean loss, loss[1..n] per sample. */
ionp_status ionp_mlp_softmax_xent(const float *logits, const int32_t *labels, const int32_t *head_off,
                                  int32_t heads, int32_t n, int32_t total, float *dlogits, float *loss,
                                  void *stream);
/* The same loss against soft targets: targets[n, total] holds non-negative
 * weights over each head's values (a distribution scaled by the sample's
 * weight m: a head's targets sum to m); loss = -sum_h sum_i q_i
 * log softmax(z)_i (over n samples, averaged), dlogits = (m softmax(z) - q)
 * / n. */
ionp_status ionp_mlp_softmax_xent_soft(const float *logits, const float *targets, const int32_t *head_off,
                                       int32_t heads, int32_t n, int32_t total, float *dlogits, float *loss,
                                       void *stream);
/* Adam step over `count` parameters. */
ionp_status ionp_mlp_adam(float *param, const float *grad, float *m, float *v, int64_t count, float lr,
                          float beta1, float beta2, float eps, int32_t step, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* IONP_HIP_H */
