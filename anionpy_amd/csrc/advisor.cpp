// advisor.cpp -- the learned launch configuration inside the library.  No
// counterpart in the reference (its CPU loops have no launch knobs).
//
// Calls made without a configuration (cfg == NULL) take the one the cost
// model proposes for their layout.  Proposals are cached by the layout's
// feature vector (ionp_call_features).  A layout seen for the first time runs
// with the built-in heuristic while its features wait in a queue; the queue
// is evaluated by the MLP's HIP forward kernels on the caller's stream (one
// batch of up to kMaxBatch layouts in flight per device), the logits come
// back by an asynchronous copy into pinned host memory, and a later call
// installs them once hipEventQuery reports the copy done.  No call waits for
// the model; a call on a stream under graph capture only reads the cache
// (no event query, launch or timing enters the capture).
//
// A proposal is then verified on the calls themselves: the layout's early
// calls (run with the heuristic while the proposal is pending; not the very
// first, which pays one-time costs) and the proposal's first calls are timed
// with event pairs, three of each, and the layout keeps whichever ran faster
// (a trial).  The events are read with
// hipEventQuery by later calls, so verifying waits for nothing either.
//
// Whatever the model proposes, results are bit-identical: a configuration
// only moves work between threads.
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "advisor.h"
#include "planner.h"

namespace ionp {
namespace {

constexpr int kMaxBatch = 256;      // layouts per forward pass
constexpr size_t kMaxQueue = 4096;  // layouts waiting for a pass (later ones: heuristic until queued)
constexpr int kTrialSamples = 3;    // timed calls per arm (heuristic, proposal) of a trial
constexpr int kTrialAttempts = 8;   // timings issued per arm before deciding on what was measured
constexpr size_t kMaxTrials = 1024; // layouts under trial at once (later ones keep the proposal untried)
constexpr size_t kMaxCache = 1u << 20;  // layouts with a proposal (~150 B each); later new layouts: heuristic

struct Key {
    float f[IONP_NFEATURES];
    bool operator==(const Key &o) const { return std::memcmp(f, o.f, sizeof(f)) == 0; }
};
struct KeyHash {
    size_t operator()(const Key &k) const {
        const unsigned char *p = reinterpret_cast<const unsigned char *>(k.f);
        uint64_t h = 1469598103934665603ull;  // FNV-1a
        for (size_t i = 0; i < sizeof(k.f); i++) h = (h ^ p[i]) * 1099511628211ull;
        return (size_t)h;
    }
};

struct Model {
    std::vector<int> sizes;     // nin, hidden..., nout
    std::vector<float> params;  // per layer: W (in x out, row-major) then b (out)
    std::vector<int> heads;     // values per knob: block, vec, blocks_per_cu, split (, fallback)
    std::vector<int> values;    // the knob values, head after head
    int maxw = 0;
};

struct Device {
    bool ready = false, broken = false;
    float *params = nullptr, *x = nullptr, *act[2] = {nullptr, nullptr};
    float *h_x = nullptr, *h_logits = nullptr;  // pinned
    hipEvent_t done = nullptr;
    std::vector<Key> inflight;
    std::vector<hipEvent_t> spare;  // timing events of finished trials, for reuse
};

// a layout's proposal: `decided` once verified (or not worth verifying);
// a decided zero configuration means the heuristic
struct Entry {
    ionp_config cfg;
    bool decided;
};

// timing slots of a trial: free -> reserved (a call will record the pair)
// -> recorded (both events recorded; read back by a later call) -> free
enum : uint8_t { SLOT_FREE = 0, SLOT_RESERVED = 1, SLOT_RECORDED = 2 };
enum { ARM_HEURISTIC = 0, ARM_PROPOSAL = 1 };
struct Trial {
    int dev = 0;
    int calls = 0;  // calls seen while pending: the first (cold) one is not timed
    hipEvent_t ev[2][kTrialSamples][2] = {};
    uint8_t state[2][kTrialSamples] = {};
    int valid[2] = {0, 0}, attempts[2] = {0, 0};
    float best[2] = {INFINITY, INFINITY};
};

std::mutex g_mu;
std::atomic<bool> g_on{false};
Model *g_model = nullptr;
std::unordered_map<Key, Entry, KeyHash> g_cache;
std::unordered_map<Key, char, KeyHash> g_pending;  // queued or in flight
std::unordered_map<Key, Trial *, KeyHash> g_trials;
std::vector<Key> g_queue;
std::vector<Device> g_dev;  // by device ordinal
int64_t g_hits = 0, g_heuristic = 0, g_inferred = 0, g_kept = 0, g_reverted = 0;
uint64_t g_generation = 0;  // bumped by ionp_advisor_set: older timings are void

void release(Device &D, int ordinal) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(ordinal);
    if (D.done) {
        (void)hipEventSynchronize(D.done);
        (void)hipEventDestroy(D.done);
    }
    for (float *p : {D.params, D.x, D.act[0], D.act[1]})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : D.spare) (void)hipEventDestroy(e);
    if (D.h_x) (void)hipHostFree(D.h_x);
    if (D.h_logits) (void)hipHostFree(D.h_logits);
    (void)hipSetDevice(cur);
    D = Device();
}

// device buffers and the weights, on the current device (first pass there)
bool setup(Device &D) {
    const Model &M = *g_model;
    const size_t nin = M.sizes.front(), nout = M.sizes.back();
    bool ok = hipMalloc((void **)&D.params, M.params.size() * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&D.x, kMaxBatch * nin * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&D.act[0], kMaxBatch * (size_t)M.maxw * sizeof(float)) == hipSuccess &&
              hipMalloc((void **)&D.act[1], kMaxBatch * (size_t)M.maxw * sizeof(float)) == hipSuccess &&
              hipHostMalloc((void **)&D.h_x, kMaxBatch * nin * sizeof(float), hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc((void **)&D.h_logits, kMaxBatch * nout * sizeof(float), hipHostMallocDefault) ==
                  hipSuccess &&
              hipEventCreateWithFlags(&D.done, hipEventDisableTiming) == hipSuccess &&
              hipMemcpy(D.params, M.params.data(), M.params.size() * sizeof(float), hipMemcpyHostToDevice) ==
                  hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        D.broken = true;  // this device keeps the heuristic
        return false;
    }
    D.ready = true;
    return true;
}

// per head the value of the largest logit; a fifth (fallback) head that
// picks 1 means the built-in heuristic (a zero configuration)
ionp_config decode(const float *z) {
    const Model &M = *g_model;
    int pick[5] = {0, 0, 0, 0, 0};
    int off = 0;
    for (size_t h = 0; h < M.heads.size(); h++) {
        int best = 0;
        for (int i = 1; i < M.heads[h]; i++)
            if (z[off + i] > z[off + best]) best = i;
        pick[h] = M.values[off + best];
        off += M.heads[h];
    }
    ionp_config c;
    if (M.heads.size() == 5 && pick[4] == 1) {
        std::memset(&c, 0, sizeof(c));
        return c;
    }
    c.block = pick[0];
    c.vec = pick[1];
    c.blocks_per_cu = pick[2];
    c.split = pick[3];
    return c;
}

// install a finished pass; `wait`: block until it is finished
void poll(Device &D, bool wait) {
    if (D.inflight.empty()) return;
    hipError_t e = wait ? hipEventSynchronize(D.done) : hipEventQuery(D.done);
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();  // not an error: the pass is still running
        return;
    }
    const int nout = g_model->sizes.back();
    for (size_t i = 0; i < D.inflight.size(); i++) {
        if (e == hipSuccess) {
            const ionp_config c = decode(D.h_logits + i * nout);
            // a proposal of the heuristic itself needs no trial
            const bool heuristic = !c.block && !c.vec && !c.blocks_per_cu && !c.split;
            g_cache[D.inflight[i]] = Entry{c, heuristic};
        }
        g_pending.erase(D.inflight[i]);
    }
    if (e == hipSuccess) {
        g_inferred += (int64_t)D.inflight.size();
    } else {  // the pass failed: this device keeps the heuristic
        (void)hipGetLastError();
        D.broken = true;
    }
    D.inflight.clear();
}

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return true;  // unknown: do not launch
    }
    return cs != hipStreamCaptureStatusNone;
}

// one forward pass over the front of the queue on stream s
void launch(Device &D, hipStream_t s) {
    if (g_queue.empty() || !D.inflight.empty() || D.broken || capturing(s)) return;
    if (!D.ready && !setup(D)) return;
    const Model &M = *g_model;
    const int nin = M.sizes.front(), nout = M.sizes.back();
    const int n = (int)std::min<size_t>(g_queue.size(), kMaxBatch);
    for (int i = 0; i < n; i++) std::memcpy(D.h_x + (size_t)i * nin, g_queue[i].f, nin * sizeof(float));
    D.inflight.assign(g_queue.begin(), g_queue.begin() + n);
    g_queue.erase(g_queue.begin(), g_queue.begin() + n);
    bool ok = hipMemcpyAsync(D.x, D.h_x, (size_t)n * nin * sizeof(float), hipMemcpyHostToDevice, s) == hipSuccess;
    const float *in = D.x;
    const float *p = D.params;
    const int L = (int)M.sizes.size() - 1;
    for (int l = 0; l < L && ok; l++) {
        const int a = M.sizes[l], b = M.sizes[l + 1];
        float *out = D.act[l & 1];
        ok = ionp_mlp_linear_fwd(in, p, p + (size_t)a * b, out, n, a, b, l < L - 1, s) == IONP_OK;
        p += (size_t)a * b + b;
        in = out;
    }
    ok = ok && hipMemcpyAsync(D.h_logits, in, (size_t)n * nout * sizeof(float), hipMemcpyDeviceToHost, s) == hipSuccess;
    ok = ok && hipEventRecord(D.done, s) == hipSuccess;
    if (!ok) {  // these layouts keep the heuristic
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        for (const Key &k : D.inflight) g_pending.erase(k);
        D.inflight.clear();
        D.broken = true;
    }
}

Device &device() {
    int d = 0;
    (void)hipGetDevice(&d);
    if ((size_t)d >= g_dev.size()) g_dev.resize(d + 1);
    return g_dev[d];
}

void enqueue(const Key &k) {
    if (g_cache.count(k) || g_pending.count(k) || g_queue.size() >= kMaxQueue ||
        g_cache.size() + g_pending.size() >= kMaxCache)
        return;
    g_queue.push_back(k);
    g_pending[k] = 1;
}

// ---- trials ------------------------------------------------------------------

// the layout's trial on this device, created when `create` (nullptr: none)
Trial *trial_get(const Key &k, Device &D, int dev, bool create) {
    auto it = g_trials.find(k);
    if (it != g_trials.end()) return it->second->dev == dev ? it->second : nullptr;
    if (!create || g_trials.size() >= kMaxTrials) return nullptr;
    Trial *T = new Trial();
    T->dev = dev;
    for (int a = 0; a < 2; a++)
        for (int j = 0; j < kTrialSamples; j++)
            for (int e = 0; e < 2; e++) {
                if (!D.spare.empty()) {
                    T->ev[a][j][e] = D.spare.back();
                    D.spare.pop_back();
                } else if (hipEventCreate(&T->ev[a][j][e]) != hipSuccess) {
                    (void)hipGetLastError();
                    T->ev[a][j][e] = nullptr;
                }
            }
    for (int a = 0; a < 2; a++)
        for (int j = 0; j < kTrialSamples; j++)
            for (int e = 0; e < 2; e++)
                if (!T->ev[a][j][e]) {  // no events: no trial
                    for (int a2 = 0; a2 < 2; a2++)
                        for (int j2 = 0; j2 < kTrialSamples; j2++)
                            for (int e2 = 0; e2 < 2; e2++)
                                if (T->ev[a2][j2][e2]) D.spare.push_back(T->ev[a2][j2][e2]);
                    delete T;
                    return nullptr;
                }
    g_trials[k] = T;
    return T;
}

// read back recorded timings that have finished (no wait)
void trial_read(Trial &T) {
    for (int a = 0; a < 2; a++)
        for (int j = 0; j < kTrialSamples; j++) {
            if (T.state[a][j] != SLOT_RECORDED) continue;
            const hipError_t e = hipEventQuery(T.ev[a][j][1]);
            if (e == hipErrorNotReady) {
                (void)hipGetLastError();
                continue;
            }
            float ms = 0.0f;
            if (e == hipSuccess && hipEventElapsedTime(&ms, T.ev[a][j][0], T.ev[a][j][1]) == hipSuccess) {
                T.valid[a]++;
                if (ms < T.best[a]) T.best[a] = ms;
            } else {
                (void)hipGetLastError();  // a pair that was not recorded: no sample
            }
            T.state[a][j] = SLOT_FREE;
        }
}

bool trial_busy(const Trial &T) {
    for (int a = 0; a < 2; a++)
        for (int j = 0; j < kTrialSamples; j++)
            if (T.state[a][j] != SLOT_FREE) return true;
    return false;
}

bool arm_done(const Trial &T, int a) { return T.valid[a] >= kTrialSamples || T.attempts[a] >= kTrialAttempts; }

// reserve a slot of arm `a` for this call
bool trial_reserve(Trial &T, int a, hipStream_t s, AdviseTiming *t) {
    if (arm_done(T, a)) return false;
    for (int j = 0; j < kTrialSamples; j++)
        if (T.state[a][j] == SLOT_FREE) {
            T.state[a][j] = SLOT_RESERVED;
            T.attempts[a]++;
            t->trial = &T;
            t->generation = g_generation;
            t->arm = a;
            t->slot = j;
            t->start = T.ev[a][j][0];
            t->end = T.ev[a][j][1];
            t->stream = s;
            return true;
        }
    return false;
}

// retire a trial no call is using: its events go back to the device
void trial_drop(const Key &k, Trial *T, Device &D) {
    for (int a = 0; a < 2; a++)
        for (int j = 0; j < kTrialSamples; j++)
            for (int e = 0; e < 2; e++) D.spare.push_back(T->ev[a][j][e]);
    delete T;
    g_trials.erase(k);
}

// decide from the measured arms: the heuristic replaces the proposal only
// when both were measured and the heuristic ran faster
void trial_finish(const Key &k, Trial *T, Entry &E, Device &D) {
    if (T->valid[ARM_HEURISTIC] > 0 && T->valid[ARM_PROPOSAL] > 0 && T->best[ARM_HEURISTIC] < T->best[ARM_PROPOSAL]) {
        std::memset(&E.cfg, 0, sizeof(E.cfg));
        g_reverted++;
    } else {
        g_kept++;
    }
    E.decided = true;
    trial_drop(k, T, D);
}

}  // namespace

void AdviseTiming::begin() {
    if (start) started = hipEventRecord(start, stream) == hipSuccess;
    if (start && !started) (void)hipGetLastError();
}

AdviseTiming::~AdviseTiming() {
    if (!trial) return;
    std::lock_guard<std::mutex> lock(g_mu);
    if (generation != g_generation) {
        // the trial went with the model it was for; ionp_advisor_set left
        // this reserved slot's events to the call that holds them
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(end);
        return;
    }
    bool recorded = false;
    if (started) {
        recorded = hipEventRecord(end, stream) == hipSuccess;
        if (!recorded) (void)hipGetLastError();
    }
    static_cast<Trial *>(trial)->state[arm][slot] = recorded ? SLOT_RECORDED : SLOT_FREE;
}

const ionp_config *advise(int kind, int op, int nin, const ionp_array_desc *const *ins, const ionp_array_desc *out,
                          const int32_t *axes, int32_t naxes, int32_t keepdims, hipStream_t s, ionp_config *slot,
                          AdviseTiming *timing) {
    if (!g_on.load(std::memory_order_acquire)) return nullptr;
    ionp_call c;
    std::memset(&c, 0, sizeof(c));
    c.kind = kind;
    c.op = op;
    c.naxes = naxes;
    c.keepdims = keepdims;
    for (int i = 0; axes && i < naxes && i < IONP_MAX_DIMS; i++) c.axes[i] = axes[i];
    for (int i = 0; i < nin; i++) c.in[i] = *ins[i];
    c.out = *out;
    Key k;
    // a layout the features reject is one the call itself rejects (same
    // descriptor and axis checks), with its own error text
    if (ionp_call_features(&c, k.f) != IONP_OK) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_model) return nullptr;
    if (capturing(s)) {
        // under graph capture only the cache is read: no event queries, no
        // passes, no timing (a capture in global mode forbids them); a new
        // layout is queued for the first call outside the capture
        auto it = g_cache.find(k);
        if (it == g_cache.end()) {
            g_heuristic++;
            enqueue(k);
            return nullptr;
        }
        g_hits++;
        *slot = it->second.cfg;
        return slot;
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    Device &D = device();
    poll(D, false);
    auto it = g_cache.find(k);
    if (it == g_cache.end()) {  // proposal pending: the heuristic, timed for the trial
        g_heuristic++;
        enqueue(k);
        launch(D, s);
        if (Trial *T = trial_get(k, D, dev, true)) {
            trial_read(*T);
            // a layout's first call pays one-time costs (first touch of its
            // buffers, code loading): not a sample
            if (T->calls++ > 0) trial_reserve(*T, ARM_HEURISTIC, s, timing);
        }
        return nullptr;
    }
    Entry &E = it->second;
    auto tr = g_trials.find(k);
    if (E.decided && tr != g_trials.end() && tr->second->dev == dev) {  // leftover heuristic timings
        trial_read(*tr->second);
        if (!trial_busy(*tr->second)) trial_drop(k, tr->second, D);
    } else if (!E.decided) {
        Trial *T = trial_get(k, D, dev, true);
        if (!T) {
            if (g_trials.size() >= kMaxTrials) E.decided = true;  // no room: the proposal untried
        } else {
            trial_read(*T);
            if (arm_done(*T, ARM_HEURISTIC) && arm_done(*T, ARM_PROPOSAL) && !trial_busy(*T)) {
                trial_finish(k, T, E, D);
            } else if (!trial_reserve(*T, ARM_PROPOSAL, s, timing) && trial_reserve(*T, ARM_HEURISTIC, s, timing)) {
                g_heuristic++;
                return nullptr;
            }
        }
    }
    g_hits++;
    *slot = E.cfg;
    return slot;
}

}  // namespace ionp

using namespace ionp;

extern "C" {

ionp_status ionp_advisor_set(const int32_t *sizes, int32_t nlayers, const float *params, const int32_t *heads,
                             int32_t nheads, const int32_t *values) {
    Model *m = nullptr;
    if (nlayers > 0) {
        if (!sizes || !params || !heads || !values || (nheads != 4 && nheads != 5))
            return set_error(IONP_ERR_VALUE, "ionp_advisor_set: a model needs sizes, params and 4 knob heads "
                                             "(and optionally the fallback head)");
        m = new Model();
        m->sizes.assign(sizes, sizes + nlayers + 1);
        m->heads.assign(heads, heads + nheads);
        int nout = 0;
        for (int h : m->heads) {
            if (h <= 0) {
                delete m;
                return set_error(IONP_ERR_VALUE, "ionp_advisor_set: empty knob head");
            }
            nout += h;
        }
        size_t count = 0;
        for (int l = 0; l < nlayers; l++) {
            if (m->sizes[l] <= 0 || m->sizes[l + 1] <= 0) {
                delete m;
                return set_error(IONP_ERR_VALUE, "ionp_advisor_set: bad layer width");
            }
            count += (size_t)m->sizes[l] * m->sizes[l + 1] + m->sizes[l + 1];
            m->maxw = std::max(m->maxw, m->sizes[l + 1]);
        }
        if (m->sizes.front() != IONP_NFEATURES || m->sizes.back() != nout) {
            delete m;
            return set_error(IONP_ERR_VALUE, "ionp_advisor_set: the model must map IONP_NFEATURES features to "
                                             "the knob heads' logits");
        }
        m->values.assign(values, values + nout);
        m->params.assign(params, params + count);
    }
    std::lock_guard<std::mutex> lock(g_mu);
    g_on.store(false, std::memory_order_release);
    g_generation++;
    for (auto &kv : g_trials) {
        for (int a = 0; a < 2; a++)
            for (int j = 0; j < kTrialSamples; j++) {
                // a reserved slot's call (on another thread) still records
                // into its events: it destroys them when it sees the new
                // generation
                if (kv.second->state[a][j] == SLOT_RESERVED) continue;
                for (int e = 0; e < 2; e++) (void)hipEventDestroy(kv.second->ev[a][j][e]);
            }
        delete kv.second;
    }
    g_trials.clear();
    for (size_t d = 0; d < g_dev.size(); d++) release(g_dev[d], (int)d);
    g_dev.clear();
    delete g_model;
    g_model = m;
    g_cache.clear();
    g_pending.clear();
    g_queue.clear();
    g_hits = g_heuristic = g_inferred = g_kept = g_reverted = 0;
    g_on.store(m != nullptr, std::memory_order_release);
    return IONP_OK;
}

ionp_status ionp_advisor_prime(const ionp_call *calls, int32_t n, void *stream) {
    if (n < 0 || (n > 0 && !calls)) return set_error(IONP_ERR_VALUE, "ionp_advisor_prime: bad call list");
    std::vector<Key> keys;
    for (int32_t i = 0; i < n; i++) {
        Key k;
        if (ionp_status s = ionp_call_features(&calls[i], k.f)) return s;
        keys.push_back(k);
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_model) return IONP_OK;
    hipStream_t s = (hipStream_t)stream;
    if (capturing(s)) return set_error(IONP_ERR_VALUE, "ionp_advisor_prime: the stream is capturing a graph");
    Device &D = device();
    for (const Key &k : keys) enqueue(k);
    while (!D.inflight.empty() || !g_queue.empty()) {
        poll(D, true);
        launch(D, s);
        if (D.broken) break;
    }
    return D.broken ? set_error(IONP_ERR_DEVICE, "ionp_advisor_prime: the model could not run on this device")
                    : IONP_OK;
}

ionp_status ionp_advisor_stats(int64_t *stats) {
    if (!stats) return set_error(IONP_ERR_VALUE, "null argument");
    std::lock_guard<std::mutex> lock(g_mu);
    stats[0] = g_hits;
    stats[1] = g_heuristic;
    stats[2] = g_inferred;
    stats[3] = (int64_t)g_pending.size();
    stats[4] = g_kept;
    stats[5] = g_reverted;
    stats[6] = (int64_t)g_trials.size();
    return IONP_OK;
}

ionp_status ionp_advisor_lookup(const ionp_call *call, ionp_config *cfg, int32_t *found) {
    if (!call || !cfg || !found) return set_error(IONP_ERR_VALUE, "null argument");
    Key k;
    if (ionp_status s = ionp_call_features(call, k.f)) return s;
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_cache.find(k);
    *found = it != g_cache.end();
    if (*found) *cfg = it->second.cfg;
    return IONP_OK;
}

}  // extern "C"
