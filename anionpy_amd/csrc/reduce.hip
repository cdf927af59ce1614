// reduce.hip -- gfx950 reduction kernels: add / maximum / minimum .reduce
// along any axes and argmax / argmin.
//
// Replaces the reference's reduce engine: reduce_axis_pairwise_narrow_wide
// (ionp-core/src/ufunc.rs:6658-6799) with pairwise_group / pairwise_sum /
// tree_combine (6371-6531), sequential_group (6546) and reduce_axis_argext
// (8037-8111).
//
// Float add.reduce must equal NumPy's result bit for bit.  NumPy's buffered
// iterator hands its add loop one window of at most `bufsize` (8192)
// elements of the innermost reduced block W at a time, and the loop does
// acc = acc + pairwise(window).  So:
//   pw_windows  - one workgroup per window evaluates NumPy's exact pairwise
//                 tree (8-accumulator 128-element leaves, split points
//                 rounded to multiples of 8) from an LDS copy of the window;
//   pw_fold     - the ordered fold acc = 0; acc += partial[i] per output;
//   pw_narrow   - W empty (reduced axes outside a kept axis): one thread per
//                 output folds its elements in iterator order, with loads
//                 batched ahead of the dependent add chain.
// maximum/minimum: a window's effect on the running value is summarised
// (best value, tie flag, NaN outcome) in parallel and the summaries are
// combined with an associative, order-preserving rule that reproduces the
// x86 AVX-512 reduce loop NumPy runs (lane order, canonical NaN).
// Integer reductions and argmax/argmin are order-free (associative).
#include "reduce.h"

#include <atomic>
#include <mutex>

#include "elementwise.h"
#include "device_ops.h"

namespace ionp {

// ---------------------------------------------------------------------------
// scratch
// ---------------------------------------------------------------------------
static std::mutex g_scratch_mu;
static void *g_scratch = nullptr;
static size_t g_scratch_bytes = 0;
static void *g_stage = nullptr;
static size_t g_stage_bytes = 0;

// Caller-owned scratch regions installed on this thread (ScratchScope), and
// the recorder of the largest requests (ScratchRecord).
static thread_local ScratchArena *t_arena = nullptr;
static thread_local ScratchUse *t_record = nullptr;

ScratchScope::ScratchScope(ScratchArena *a) : old(t_arena) { t_arena = a; }
ScratchScope::~ScratchScope() { t_arena = old; }
ScratchRecord::ScratchRecord(ScratchUse *u) : old(t_record) { t_record = u; }
ScratchRecord::~ScratchRecord() { t_record = old; }

// a request against an installed arena: its region, or null (and the
// arena marked short) when the region is smaller
static void *arena_take(char *base, size_t have, size_t bytes, size_t ScratchUse::*need) {
    ScratchArena &A = *t_arena;
    if (bytes > A.need.*need) A.need.*need = bytes;
    if (bytes > have || !base) {
        A.short_ = true;
        return nullptr;
    }
    return base;
}

// a second grow-only buffer: inputs converted to another dtype before a
// kernel that may itself take scratch_get memory
void *stage_get(size_t bytes, hipStream_t s) {
    if (t_record && bytes > t_record->stage) t_record->stage = bytes;
    if (t_arena) return arena_take(t_arena->stage, t_arena->size.stage, bytes, &ScratchUse::stage);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (bytes <= g_stage_bytes) return g_stage;
    if (g_stage) {
        (void)hipDeviceSynchronize();
        (void)hipFree(g_stage);
        g_stage = nullptr;
        g_stage_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (hipMalloc(&g_stage, want) != hipSuccess) {
        g_stage = nullptr;
        return nullptr;
    }
    g_stage_bytes = want;
    (void)s;
    return g_stage;
}

void *scratch_get(size_t bytes, hipStream_t s) {
    if (t_record && bytes > t_record->scratch) t_record->scratch = bytes;
    if (t_arena) return arena_take(t_arena->scratch, t_arena->size.scratch, bytes, &ScratchUse::scratch);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (bytes <= g_scratch_bytes) return g_scratch;
    if (g_scratch) {
        (void)hipDeviceSynchronize();  // earlier launches may still read it
        (void)hipFree(g_scratch);
        g_scratch = nullptr;
        g_scratch_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (hipMalloc(&g_scratch, want) != hipSuccess) {
        g_scratch = nullptr;
        return nullptr;
    }
    g_scratch_bytes = want;
    (void)s;
    return g_scratch;
}

// One workgroup per window g = (o, r, w), windows enumerated output-major.
template <typename T>
// tail_nwin > 0: only the last (partial) window of each output, written to
// partial index o * tail_nwin + tail_nwin - 1 (the full windows go to pw_stream).
__global__ __launch_bounds__(kPwThreads) void pw_windows(const RedArgs A, int64_t nwindows, int direct, int64_t tail_nwin) {
    __shared__ T xs[kPwMaxWin];
    __shared__ T racc[kPwMaxWin / 8];
    __shared__ T val[kPwMaxWin / 8];
    __shared__ int leaf_lo[kPwMaxWin / 64 + 2], leaf_n[kPwMaxWin / 64 + 2];
    __shared__ int cnt[1];
    constexpr int LT = FDT<T>::v;
    for (int64_t g = blockIdx.x; g < nwindows; g += gridDim.x) {
        const int64_t per_out = A.rsize * A.nwin;
        int64_t o, r, w;
        if (tail_nwin) {
            o = g;
            r = 0;
            w = tail_nwin - 1;
        } else {
            o = g / per_out;
            const int64_t rem = g - o * per_out;
            r = rem / A.nwin;
            w = rem - r * A.nwin;
        }
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ib += r_offset(A, r);
        const int64_t s0 = w * A.bufsize;
        const int L = (int)min(A.bufsize, A.wsize - s0);
        for (int i = threadIdx.x; i < L; i += blockDim.x) xs[i] = ld<LT>(A.in_dt, A.in, ib + w_offset(A, s0 + i));
        __syncthreads();
        const T res = pairwise_block<T>([&](int i) { return xs[i]; }, L, racc, val, leaf_lo, leaf_n, cnt);
        if (threadIdx.x == 0) {
            if (direct) ((T *)A.out)[ob] = fadd(fseed<T>(A), res);
            else ((T *)A.scratch)[tail_nwin ? o * tail_nwin + w : g] = res;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Full 8192-element windows of contiguous rows: NumPy's tree on a balanced
// 64-leaf shape, evaluated from registers (pw_stream), with the ordered fold
// of window partials running inside the same launch (pw_fold_ring).
// ---------------------------------------------------------------------------
// Leaf l of a window holds elements [128 l, 128 l + 128); its accumulator j
// sums elements 8k + j (k = 0..15) in order, and the window value is the
// balanced tree over the 64 leaf values (ufunc.rs:6416-6448 with n = 8192:
// every split is exact, every leaf has 128 elements).
//
// Producer wavefronts take windows round-robin (window g = producer + k *
// producers, o = g / nfull, w = g % nfull), so windows finish roughly in
// order.  Each publishes partial[o][w] with one atomic store; a partial slot
// starts as a signalling-NaN sentinel that no IEEE add can produce (every
// NaN an add returns is quiet), so a slot that is not the sentinel holds its
// final value and no flag, fence or counter is needed.  The fold
// acc = 0; acc = acc + partial[w] (w order) is done
//   - for few outputs: by a folder workgroup per output (blocks 0..nout-1)
//     that polls the partial slots while the producers stream, so the
//     inherently serial fold (32768 adds for 2^28 elements) overlaps the
//     HBM traffic instead of following it;
//   - otherwise: by pw_fold after the launch.
// Producers never wait; a folder waits only for slots producers always fill
// and gives up after a bounded number of polls of one slot (a producer that
// can not be scheduled, e.g. CUs held by another process): the output is
// then NaN, `dirty` asks the next folded call to refill every slot first
// (late producers may still write some), and `err` -- host-visible memory
// -- lets the host report the failure without a synchronisation.
struct PwStreamState {
    void *partials;    // nout * nwin of T, pre-filled with the sentinel
    int64_t nfull, nwin;
    int32_t folders;   // folder workgroups (0 or nout)
    uint32_t poll_limit;  // polls of one unfilled slot before the folder gives up
    int64_t stall;     // test hook: a window index the folder never sees filled (-1: none)
    uint32_t *dirty;   // device word, or null (slots refilled on every use anyway)
    uint32_t *err;     // fine-grained host word, or null
};

template <typename T> struct Sentinel;
template <> struct Sentinel<double> { static constexpr uint64_t bits = 0x7ff0dead0000beefull; };
template <> struct Sentinel<float> { static constexpr uint64_t bits = 0x7f80beefu; };

template <typename T> __device__ inline uint64_t ld_bits(const T *p) {
    if constexpr (sizeof(T) == 8) return __hip_atomic_load((uint64_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return __hip_atomic_load((uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ inline void st_value(T *p, T v) {
    if constexpr (sizeof(T) == 8)
        __hip_atomic_store((uint64_t *)p, __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        __hip_atomic_store((uint32_t *)p, __builtin_bit_cast(uint32_t, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ inline T from_bits(uint64_t u) {
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, u);
    else return __builtin_bit_cast(T, (uint32_t)u);
}
// pairwise combine across lanes at distance m: the lower lane's value is
// the left operand, so both lanes compute left + right
template <typename T> __device__ inline T xor_combine(T v, int m, int lane) {
    const T o = __shfl_xor(v, m, 64);
    return (lane & m) ? o + v : v + o;
}

// One wavefront evaluates a whole 8192-element window (NumPy's tree with 64
// leaves of 128; pw_stream uses it to redo a window whose plain-IEEE value
// came out NaN, with EXACT set).  Every
// lane owns CH = 16 / sizeof(T) consecutive accumulators of one leaf and
// reads them with one 16-byte load per step k (k = 0..15, element
// 128 leaf + 8 k + j), so 8 / CH lanes cover a leaf, a pass covers 64 * CH / 8
// leaves and NP passes cover the window's 64 leaves.  Leaf values are
// gathered one per lane and combined by the balanced tree over lanes.
// EXACT: every add with fadd's x86 NaN rules (plain IEEE adds otherwise;
// the two agree unless a NaN arises).
template <typename T, bool EXACT> __device__ inline T pw_window_wave(const T *x, int lane) {
    constexpr int CH = 16 / sizeof(T);  // accumulators per lane
    constexpr int LPL = 8 / CH;         // lanes per leaf
    constexpr int LPW = 64 / LPL;       // leaves per pass
    constexpr int NP = 64 / LPW;        // passes per window
    using V = T __attribute__((ext_vector_type(CH)));
    auto add = [](T a, T b) { return EXACT ? fadd(a, b) : a + b; };
    auto comb = [&](T v, int m) {  // pairwise combine at lane distance m, left operand = lower lane
        const T o = __shfl_xor(v, m, 64);
        return (lane & m) ? add(o, v) : add(v, o);
    };
    const int lf = lane / LPL, jq = (lane % LPL) * CH;
    T leafv[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const T *px = x + (int64_t)(p * LPW + lf) * 128 + jq;
        V v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = *(const V *)(px + 8 * k);
        T r[CH];
#pragma unroll
        for (int c = 0; c < CH; c++) {
            r[c] = v[0][c];
#pragma unroll
            for (int k = 1; k < 16; k++) r[c] = add(r[c], v[k][c]);
        }
        T a;
        if constexpr (CH == 2) {
            a = add(r[0], r[1]);  // r0+r1 (lane holds accumulators 2m, 2m+1)
            a = comb(a, 1);       // (r0+r1)+(r2+r3)
            a = comb(a, 2);       // leaf value
        } else {
            a = add(add(r[0], r[1]), add(r[2], r[3]));
            a = comb(a, 1);  // leaf value
        }
        leafv[p] = a;
    }
    // lane L takes leaf L: pass L / LPW, held by lane (L % LPW) * LPL
    T t = leafv[0];
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const T s = __shfl(leafv[p], (lane % LPW) * LPL, 64);
        if (lane / LPW == p) t = s;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) t = comb(t, m);
    return t;  // the same in every lane
}

// Ring folder: the fold acc = acc + partial[w] of one output split over two
// wavefronts of one workgroup.  The poller (wave 1) reads the partial slots
// 1024 at a time, three blocks ahead, and copies each filled prefix into an
// LDS ring of kRing entries, then publishes the count (ctl[0]).  The chain
// (wave 0, every lane computing the same sums from broadcast LDS reads,
// issued a chunk ahead) only adds: its dependent add chain, about 11 cycles an add on MI355X, sets the
// pace (32768 windows in ~150 us), well under the producers' HBM time, where
// one wavefront that also polls and stages runs at twice that.  It returns
// its position (ctl[1]) so the poller never overwrites unread entries.
// Ordering: each wavefront's LDS operations execute in issue order, so the
// ring writes land before the count that publishes them and the chain's
// reads follow the count it read; the wavefront-scope fences keep the
// compiler from reordering across the counter accesses.  A poller that gives
// up (bounded wait on a slot producers always fill) sets ctl[2] and the
// output becomes NaN.
// The slots live in a buffer that holds the sentinel between calls: a third
// wavefront writes it back into every slot the chain has consumed, so no
// fill precedes the next call.  NaN: the chain checks acc at checkpoints
// (every 256 entries); the stretch that turned it NaN is redone from the
// ring with x86 NaN rules (fadd), and that NaN is final (an x86 add keeps a
// NaN first operand), so the partials are never needed again.
constexpr int kRing = 4096;
template <typename T>
__device__ void pw_fold_ring(const RedArgs &A, const PwStreamState &F, int64_t o, T *ring, uint32_t *ctl, int wave) {
    constexpr int Q = 16, BLK = 64 * Q;
    const int lane = threadIdx.x & 63;
    const uint32_t nwin = (uint32_t)F.nwin;
    T *slots = (T *)F.partials + o * F.nwin;
    const T *part = slots;
    auto lds_load = [](uint32_t *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
    auto lds_store = [](uint32_t *p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
    if (wave == 1) {
        __builtin_amdgcn_s_setprio(2);
        uint32_t cur = 0, idle = 0;
        bool failed = false;
        uint64_t b0[Q], b1[Q], b2[Q];
        auto poll = [&](uint64_t *b, int64_t base) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int64_t idx = base + 64 * q + lane;
                b[q] = idx < (int64_t)nwin && idx != F.stall ? ld_bits(part + idx) : Sentinel<T>::bits;
            }
        };
        poll(b0, cur);
        poll(b1, (int64_t)cur + BLK);
        poll(b2, (int64_t)cur + 2 * BLK);
        while (cur < nwin) {
            int n = 0;  // filled prefix of block b0 (stops at the first sentinel)
            bool open = true;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const int64_t idx0 = (int64_t)cur + 64 * q;
                const uint64_t ok = __ballot(b0[q] != Sentinel<T>::bits || idx0 + lane >= nwin);
                if (open) {
                    const int m = ~ok == 0 ? 64 : __builtin_ctzll(~ok);
                    const int valid = (int)min<int64_t>(64, (int64_t)nwin - idx0 > 0 ? (int64_t)nwin - idx0 : 0);
                    const int take = m < valid ? m : valid;
                    n += take;
                    open = take == 64;
                }
            }
            if (n > 0) {
                // room in the ring for entries cur .. cur + n - 1
                while (cur + n - lds_load(&ctl[1]) > (uint32_t)kRing) __builtin_amdgcn_s_sleep(1);
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                for (int q = 0; q < Q; q++)
                    if (64 * q + lane < n) ring[(cur + 64 * q + lane) & (kRing - 1)] = from_bits<T>(b0[q]);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                if (lane == 0) lds_store(&ctl[0], cur + n);
                cur += n;
            }
            if (n == BLK) {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    b0[q] = b1[q];
                    b1[q] = b2[q];
                }
                poll(b2, (int64_t)cur + 2 * BLK);
                idle = 0;
            } else if (cur < nwin) {
                if (n == 0) {
                    // wait on the first unfilled slot alone, then re-read the blocks
                    while ((int64_t)cur == F.stall || ld_bits(part + cur) == Sentinel<T>::bits) {
                        if (++idle > F.poll_limit) {  // producers always fill every slot; safety bound only
                            failed = true;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                    if (failed) break;
                } else {
                    idle = 0;
                }
                poll(b0, cur);
                poll(b1, (int64_t)cur + BLK);
                poll(b2, (int64_t)cur + 2 * BLK);
            }
        }
        if (failed) {
            // leave every slot of this output holding the sentinel again (a
            // producer still running may overwrite one: `dirty` has the next
            // folded call refill them all first)
            for (uint32_t i = cur + lane; i < nwin; i += 64) st_value(slots + i, from_bits<T>(Sentinel<T>::bits));
            if (lane == 0) {
                lds_store(&ctl[2], 1u);
                if (F.dirty) __hip_atomic_store(F.dirty, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (F.err) __hip_atomic_store(F.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        return;
    }
    if (wave == 2) {
        // the resetter: the sentinel back into every slot the chain has
        // consumed (ctl[1]), until the chain has finished (ctl[3]); its
        // stores are kept out of the poller, whose loads would wait on them
        uint32_t done = 0;
        for (;;) {
            const uint32_t fin = lds_load(&ctl[3]);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t c = __builtin_amdgcn_readfirstlane(lds_load(&ctl[1]));
            for (uint32_t j = done + lane; j < c; j += 64) st_value(slots + j, from_bits<T>(Sentinel<T>::bits));
            const bool moved = c > done;
            done = c > done ? c : done;
            if (done >= nwin || fin) break;  // fin was read before c: c is final
            if (!moved) __builtin_amdgcn_s_sleep(2);
        }
        return;
    }
    // wave 0: the chain
    __builtin_amdgcn_s_setprio(3);
    T acc = fseed<T>(A), ck_acc = acc, nan_val = acc;
    uint32_t i = 0, ck = 0;  // checkpoint: acc was ck_acc before entry ck
    bool failed = false, nan = false;  // nan: the result is nan_val
    // Plain adds; at a checkpoint, a NaN that arose since the last one is
    // redone from the ring with x86 NaN rules (fadd) and kept as the result
    // (an x86 add keeps a NaN first operand).  Entries from ck on are still
    // in the ring: the poller overwrites only below ctl[1] + kRing, and ctl[1]
    // moves only here.
    auto checkpoint = [&]() {
        if (!nan && acc != acc) {
            T r = ck_acc;
            for (uint32_t j = ck; j < i; j++) {
                r = fadd(r, ring[j & (kRing - 1)]);
                if (r != r) break;
            }
            nan_val = r;
            nan = true;
        }
        ck_acc = acc;
        ck = i;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        lds_store(&ctl[1], i);
    };
    while (i < nwin) {
        // whole 16-entry chunks only, so i stays 16-aligned; the last few
        // entries once every slot is in
        const uint32_t pub = __builtin_amdgcn_readfirstlane(lds_load(&ctl[0]));  // the same in every lane
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t avail = pub == nwin ? pub : pub & ~15u;
        if (avail <= i) {
            if (lds_load(&ctl[2])) {
                failed = true;
                break;
            }
            __builtin_amdgcn_s_sleep(0);
            continue;
        }
        // 16-entry chunks (16-aligned: no wrap inside one), two per step in
        // alternating registers; the next chunk's LDS reads are issued before
        // the current chunk's adds and the scheduling barriers keep them there
        using V2 = T __attribute__((ext_vector_type(16 / sizeof(T))));
        constexpr int NV = 16 * sizeof(T) / 16;  // 16-byte reads per chunk
        auto rd = [&](V2 *c, uint32_t at) {
            const V2 *r = (const V2 *)(ring + (at & (kRing - 1)));
#pragma unroll
            for (int k = 0; k < NV; k++) c[k] = r[k];
        };
        auto addc = [&](const V2 *c) {
#pragma unroll
            for (int k = 0; k < NV; k++)
#pragma unroll
                for (int e = 0; e < 16 / (int)sizeof(T); e++) acc = acc + c[k][e];
        };
        if (avail - i >= 16) {
            V2 ca[NV], cb[NV];
            rd(ca, i);
            while (avail - i >= 32) {  // chunks i (in ca) and i + 16
                rd(cb, i + 16);
                __builtin_amdgcn_sched_barrier(0);
                addc(ca);
                __builtin_amdgcn_sched_barrier(0);
                rd(ca, i + 32);  // entries past avail are read but not used
                __builtin_amdgcn_sched_barrier(0);
                addc(cb);
                __builtin_amdgcn_sched_barrier(0);
                i += 32;
                if ((i & 255) < 32) checkpoint();
            }
            if (avail - i >= 16) {
                addc(ca);
                i += 16;
            }
        }
        for (; i < avail; i++) acc = acc + ring[i & (kRing - 1)];
        checkpoint();
    }
    if (nan) acc = nan_val;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    lds_store(&ctl[3], 1u);  // finished: ctl[1] is final
    if (lane == 0) {
        if (failed) acc = canonical_nan(T(0));
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((T *)A.out)[ob] = acc;
    }
}

// Streaming producers: every wavefront owns whole windows (window g = its
// global wave index + k * producer waves) and reads each one as batches of
// 16 KB, one fully coalesced 1 KB load per instruction (lane L takes bytes
// 16 L .. 16 L + 15 of chunk k), which HBM serves faster than loading the
// leaf layout directly (measured on MI355X: 6.6-7.0 against 5.7-6.2 TB/s for
// reads alone).  The batch is moved through the wave's own LDS region into
// the leaf layout (lane: accumulators jq .. jq+CH-1 of leaf lf, 16 steps of
// one 16-byte read), so every add keeps NumPy's order.  A batch holds LPB whole
// leaves, a complete subtree of the window's balanced tree, so its value is
// combined across lanes at once and the window's NB batch values finish the
// tree in registers: no workgroup barrier.  The next batch's loads are issued
// before the current batch is combined.  LDS rows are padded by 8 elements
// per leaf, so the leaf-layout reads hit distinct banks.  NT: nontemporal
// loads (inputs larger than the Infinity Cache).
constexpr int kPsWaves = 4;  // wavefronts per workgroup
template <typename T> constexpr int ps_leaves() { return 128 / sizeof(T); }  // leaves per batch: 16 f64, 32 f32

template <typename T, bool NT>
__global__ __launch_bounds__(64 * kPsWaves) __attribute__((amdgpu_waves_per_eu(2))) void pw_stream(const RedArgs A, const PwStreamState F, int64_t nwindows) {
    constexpr int CH = 16 / sizeof(T);  // elements per 16-byte load
    constexpr int LPL = 8 / CH;         // lanes per leaf
    constexpr int LPB = ps_leaves<T>(); // leaves per batch (16 loads of 1 KB)
    constexpr int NB = 64 / LPB;        // batches per window (4 f64, 2 f32)
    constexpr int LS = 128 + 8;         // LDS leaf stride (elements)
    using V = T __attribute__((ext_vector_type(CH)));
    __shared__ T lds[kPsWaves][LPB * LS];
    __shared__ uint32_t ctl[4];
    static_assert(kPsWaves * LPB * LS >= kRing, "the folder's ring reuses the batch regions");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if ((int64_t)blockIdx.x < F.folders) {
        if (threadIdx.x < 4) ctl[threadIdx.x] = 0;
        __syncthreads();
        if (wave < 3) pw_fold_ring<T>(A, F, blockIdx.x, &lds[0][0], ctl, wave);
        return;
    }
    T *L = lds[wave];
    const int lf = lane / LPL, jq = (lane % LPL) * CH;
    const int64_t pwaves = ((int64_t)gridDim.x - F.folders) * kPsWaves;
    auto window_base = [&](int64_t g) {
        const int64_t o = g / F.nfull, w = g - o * F.nfull;
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        return (const T *)A.in + ib + w * 8192;
    };
    auto load = [&](V *r, const T *p) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const V *q = (const V *)(p + (k * 64 + lane) * CH);
            r[k] = NT ? __builtin_nontemporal_load(q) : *q;
        }
    };
    // one batch: registers (coalesced layout) -> LDS -> leaf layout -> subtree value
    auto batch = [&](const V *r) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            const int e = (k * 64 + lane) * CH;  // element of the batch; CH elements never straddle a leaf
            *(V *)(L + e + (e >> 7) * 8) = r[k];
        }
        // the wave's DS operations execute in order; the fences keep the
        // compiler from moving the reads above the other lanes' writes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        V v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = *(const V *)(L + lf * LS + 8 * k + jq);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        V r0 = v[0];
#pragma unroll
        for (int k = 1; k < 16; k++) r0 = r0 + v[k];
        T a;
        if constexpr (CH == 2) {
            a = r0[0] + r0[1];
            a = xor_combine(a, 1, lane);
            a = xor_combine(a, 2, lane);  // leaf value
        } else {
            a = (r0[0] + r0[1]) + (r0[2] + r0[3]);
            a = xor_combine(a, 1, lane);  // leaf value
        }
#pragma unroll
        for (int m = LPL; m < 64; m <<= 1) a = xor_combine(a, m, lane);  // the batch's subtree
        return a;
    };
    int64_t g = ((int64_t)blockIdx.x - F.folders) * kPsWaves + wave;
    if (g >= nwindows) return;
    V ra[16], rb[16];
    load(ra, window_base(g));
    for (; g < nwindows; g += pwaves) {
        const T *xw = window_base(g);
        // the next window's first batch; after its last window every
        // wavefront reads the same first window, which HBM then serves about
        // once (an unconditional load keeps the wait counts static across
        // windows)
        const T *xn = g + pwaves < nwindows ? window_base(g + pwaves) : window_base(0);
        T sb[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) {
            // batch b sits in ra (b even) or rb (b odd); request the next one into the other
            V *cur = (b & 1) ? rb : ra, *nxt = (b & 1) ? ra : rb;
            load(nxt, b + 1 < NB ? xw + (b + 1) * LPB * 128 : xn);
            sb[b] = batch(cur);
        }
        T t;
        if constexpr (NB == 4) t = (sb[0] + sb[1]) + (sb[2] + sb[3]);
        else t = sb[0] + sb[1];
        if (t != t) t = pw_window_wave<T, true>(xw, lane);  // NaN: x86 NaN rules, exactly
        if (lane == 0) {
            const int64_t o = g / F.nfull, w = g - o * F.nfull;
            st_value((T *)F.partials + o * F.nwin + w, t);
        }
    }
}

template <typename T> __global__ void fill_sentinel(T *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = from_bits<T>(Sentinel<T>::bits);
}

// One workgroup: after a folder gave up (dirty set), the sentinel into all
// n slots, then dirty cleared; otherwise nothing.  Runs in stream order
// ahead of the next folded call, so every late producer has finished.
template <typename T> __global__ __launch_bounds__(256) void slots_refill(T *p, int64_t n, uint32_t *dirty) {
    if (__hip_atomic_load(dirty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = from_bits<T>(Sentinel<T>::bits);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(dirty, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}


// Ordered fold per output: acc = 0; acc = acc + partial[i] (ufunc.rs:7137
// zero seed; the outer-axis sequential fold of pairwise_group 6504-6531).
// Partial slots of the in-launch folds: a grow-only buffer that holds the
// sentinel between calls (the folders write it back into every slot they
// consume), refilled only when it grows or its element size changes.
static void *g_slots = nullptr;
static size_t g_slots_bytes = 0;
static int g_slots_isz = 0;  // element size of the sentinel it holds (0: none)
template <typename T> static T *slots_get(int64_t n, hipStream_t s) {
    const size_t bytes = (size_t)n * sizeof(T);
    if (t_record && bytes > t_record->slots) t_record->slots = bytes;
    if (t_arena) {
        // a caller's region keeps no state between uses: its n slots are
        // filled with the sentinel ahead of every use (a kernel, so a
        // captured graph refills them on every replay)
        T *p = (T *)arena_take(t_arena->slots, t_arena->size.slots, bytes, &ScratchUse::slots);
        if (p && n > 0)
            hipLaunchKernelGGL(fill_sentinel<T>, dim3((unsigned)std::min<int64_t>(1024, (n + 255) / 256)), dim3(256), 0, s,
                               p, n);
        return p;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (bytes > g_slots_bytes) {
        if (g_slots) {
            (void)hipDeviceSynchronize();  // earlier launches may still use it
            (void)hipFree(g_slots);
            g_slots = nullptr;
            g_slots_bytes = 0;
        }
        const size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        if (hipMalloc(&g_slots, want) != hipSuccess) {
            g_slots = nullptr;
            return nullptr;
        }
        g_slots_bytes = want;
        g_slots_isz = 0;
    }
    if (g_slots_isz != (int)sizeof(T)) {
        const int64_t cnt = (int64_t)(g_slots_bytes / sizeof(T));
        hipLaunchKernelGGL(fill_sentinel<T>, dim3((unsigned)std::min<int64_t>(1024, (cnt + 255) / 256)), dim3(256), 0, s,
                           (T *)g_slots, cnt);
        g_slots_isz = (int)sizeof(T);
    }
    return (T *)g_slots;
}
static void slots_dirty() {
    if (t_arena) return;  // refilled on every use
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    g_slots_isz = 0;
}

template <typename T> __global__ __launch_bounds__(256) void pw_fold(const RedArgs A) {
    const int64_t per_out = A.rsize * A.nwin;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        const T *p = (const T *)A.scratch + o * per_out;
        T acc = fseed<T>(A);
        int64_t i = 0;
        constexpr int U = 16;
        for (; i + U <= per_out; i += U) {
            T v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = p[i + u];
#pragma unroll
            for (int u = 0; u < U; u++) acc = fadd(acc, v[u]);
        }
        for (; i < per_out; i++) acc = fadd(acc, p[i]);
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((T *)A.out)[ob] = acc;
    }
}

// Windows of at most 128 elements (one pairwise leaf each) and many
// outputs: one thread per output runs acc = 0; acc = acc + pw(window r) over
// the outer reduced index r in order (the same value pw_windows + pw_fold
// produce, without a workgroup per window).
template <typename T> __global__ __launch_bounds__(256) void pw_small(const RedArgs A) {
    constexpr int LT = FDT<T>::v;
    const int L = (int)A.wsize;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        T acc = fseed<T>(A);
        for (int64_t r = 0; r < A.rsize; r++) {
            const int64_t base = ib + r_offset(A, r);
            acc = fadd(acc, pw_leaf<T>([&](int i) { return ld<LT>(A.in_dt, A.in, base + w_offset(A, i)); }, L));
        }
        ((T *)A.out)[ob] = acc;
    }
}

// W empty: per output, acc = 0; acc = acc + x for each reduced element in
// iterator order (narrow_chain); IT is the input element type (the host
// dispatches on it, so each kernel is one straight loop).
template <typename T, typename IT> __global__ __launch_bounds__(256) void pw_narrow(const RedArgs A) {
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((T *)A.out)[ob] = narrow_chain<T, IT, (sizeof(T) == 8 ? 24 : 32)>(A, ib);
    }
}

// the same with a wavefront per output (narrow_wave_ok)
template <typename T, typename IT> __global__ __launch_bounds__(256) void pw_narrow_wave(const RedArgs A) {
    // a block of 512 elements of the chain per wavefront: about 3 us of adds,
    // enough to cover the next block's HBM loads
    constexpr int B = 8;
    __shared__ __align__(16) T lb[4][64 * B];
    const int64_t w0 = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int64_t o = w0; o < A.nout; o += (int64_t)gridDim.x * (blockDim.x >> 6)) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const T acc = narrow_wave_chain<T, IT, 8, B>(A, ib, lb[threadIdx.x >> 6]);
        if ((threadIdx.x & 63) == 0) ((T *)A.out)[ob] = acc;
    }
}

template <typename T> static void launch_narrow(const RedArgs &A, int bpc, hipStream_t s);

// ---------------------------------------------------------------------------
// multiply on floats (sequential_group_ordered, ufunc.rs:6583): no pairwise
// tree; per output one running product over the group in iterator order
// (outer reduced index, then the W block), from initial= (default 1).  Plain
// IEEE multiplies agree with fmul until a NaN appears, and a NaN stays, so
// an output that ends NaN is recomputed with fmul's x86 NaN rules.
// ---------------------------------------------------------------------------
// group element q -> input offset: q = w + wsize * r
__device__ inline int64_t group_offset(const RedArgs &A, int64_t q) {
    if (A.nr == 0) return w_offset(A, q);
    const int64_t r = q / A.wsize, w = q - r * A.wsize;
    return w_offset(A, w) + r_offset(A, r);
}
// the same element's where= mask offset
__device__ inline int64_t mask_group_offset(const RedArgs &A, int64_t q) {
    if (A.nr == 0) return mask_w_offset(A, q);
    const int64_t r = q / A.wsize, w = q - r * A.wsize;
    return mask_w_offset(A, w) + mask_r_offset(A, r);
}

template <typename T> __device__ inline T pseed(const RedArgs &A) {
    if (!A.has_init) return T(1);
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, A.init_bits);
    else return __builtin_bit_cast(T, (uint32_t)A.init_bits);
}
// group element q (iterator order) -> input offset from the output's base
__device__ inline int64_t prod_offset(const RedArgs &A, int64_t q) {
    if (A.nw == 0) return A.nr == 1 ? q * A.r_st[0] : r_offset(A, q);
    if (A.nr == 0 && A.nw == 1) return q * A.w_st[0];
    return group_offset(A, q);
}
__device__ inline int64_t prod_moffset(const RedArgs &A, int64_t q) {
    if (A.nw == 0) return A.nr == 1 ? q * A.r_m_st[0] : mask_r_offset(A, q);
    if (A.nr == 0 && A.nw == 1) return q * A.w_m_st[0];
    return mask_group_offset(A, q);
}
// group element q as a factor: a where= masked-out element is 1 (x * 1 == x
// exactly, so skipping it and multiplying by 1 agree)
template <typename T, typename IT, bool MASK> __device__ inline T prod_factor(const RedArgs &A, int64_t ib, int64_t mb,
                                                                             int64_t q) {
    constexpr int LT = FDT<T>::v;
    if constexpr (MASK)
        if (!A.mask[mb + prod_moffset(A, q)]) return T(1);
    return ld_t<LT, IT>(A.in, ib + prod_offset(A, q));
}
template <typename T, typename IT, bool MASK> __device__ inline T prod_redo(const RedArgs &A, int64_t ib, int64_t mb) {
    T acc = pseed<T>(A);
    for (int64_t q = 0; q < A.group; q++) acc = fmul(acc, prod_factor<T, IT, MASK>(A, ib, mb, q));
    return acc;
}
// one thread per output, U loads in flight ahead of the multiply chain
template <typename T, typename IT, bool MASK> __global__ __launch_bounds__(256) void prod_cols(const RedArgs A) {
    constexpr int U = 16;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t mb = MASK ? mask_k_offset(A, (uint32_t)o) : 0;
        const int64_t n = A.group;
        T acc = pseed<T>(A);
        int64_t q = 0;
        for (; q + U <= n; q += U) {
            T v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = prod_factor<T, IT, MASK>(A, ib, mb, q + u);
#pragma unroll
            for (int u = 0; u < U; u++) acc = acc * v[u];
        }
        for (; q < n; q++) acc = acc * prod_factor<T, IT, MASK>(A, ib, mb, q);
        if (acc != acc) acc = prod_redo<T, IT, MASK>(A, ib, mb);
        ((T *)A.out)[ob] = acc;
    }
}
// few long groups: a wavefront per output loads blocks of 64 * B group
// elements (the next block in flight), and the chain, the same in every
// lane, reads them lane by lane in iterator order
template <typename T, typename IT, bool MASK> __global__ __launch_bounds__(256) void prod_wave(const RedArgs A) {
    constexpr int B = 4, BLK = 64 * B;
    constexpr int CH = 16 / sizeof(T), R = 8, CK = R * CH, NCK = BLK / CK;
    using V = T __attribute__((ext_vector_type(CH)));
    // each block of factors is staged in the wavefront's LDS region and the
    // chain reads it back with broadcast 16-byte reads, a chunk ahead (the
    // same scheme as narrow_wave_chain)
    __shared__ __align__(16) T lbs[4][BLK];
    T *lb = lbs[threadIdx.x >> 6];
    auto wave_sync = [] {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    const int lane = threadIdx.x & 63;
    const int64_t w0 = blockIdx.x * (int64_t)(blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int64_t o = w0; o < A.nout; o += (int64_t)gridDim.x * (blockDim.x >> 6)) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t mb = MASK ? mask_k_offset(A, (uint32_t)o) : 0;
        const int64_t n = A.group;
        T cur[B], nxt[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int64_t q = k * 64 + lane;
            cur[k] = q < n ? prod_factor<T, IT, MASK>(A, ib, mb, q) : T(1);
            nxt[k] = T(1);
        }
        T acc = pseed<T>(A);
        for (int64_t qb = 0; qb < n; qb += BLK) {
            const int64_t nb = qb + BLK;
            if (nb < n) {
#pragma unroll
                for (int k = 0; k < B; k++) {
                    const int64_t q = nb + k * 64 + lane;
                    nxt[k] = q < n ? prod_factor<T, IT, MASK>(A, ib, mb, q) : T(1);
                }
            }
#pragma unroll
            for (int k = 0; k < B; k++) lb[k * 64 + lane] = cur[k];
            wave_sync();
            if (nb <= n) {
                V ch[2][R];
#pragma unroll
                for (int j = 0; j < R; j++) ch[0][j] = *(const V *)(lb + j * CH);
#pragma unroll
                for (int c = 0; c < NCK; c++) {
                    if (c + 1 < NCK) {
#pragma unroll
                        for (int j = 0; j < R; j++) ch[(c + 1) & 1][j] = *(const V *)(lb + (c + 1) * CK + j * CH);
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int j = 0; j < R; j++)
#pragma unroll
                        for (int e = 0; e < CH; e++) acc = acc * ch[c & 1][j][e];
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
                const int cnt = (int)(n - qb);
                for (int e = 0; e < cnt; e++) acc = acc * lb[e];
            }
            wave_sync();  // every lane has read the block before the next one is stored
#pragma unroll
            for (int k = 0; k < B; k++) cur[k] = nxt[k];
        }
        if (lane == 0) ((T *)A.out)[ob] = acc != acc ? prod_redo<T, IT, MASK>(A, ib, mb) : acc;
    }
}

// ---------------------------------------------------------------------------
// maximum / minimum on floats
// ---------------------------------------------------------------------------
// Effect of a run of windows on the running value p (an associative,
// order-preserving summary):
//   p NaN      -> canonical NaN if any window ran NumPy's SIMD reduce
//                 (`simd`), else p unchanged;
//   otherwise  -> `nanv` if the run produces a NaN, else the best value M
//                 if it beats p, or ties p with `flag` set (its element sits
//                 after every copy of p in the loop's evaluation order).
template <typename T> struct ExtSum {
    T m, nanv;
    uint32_t has_m, flag, has_nan, simd;
};

template <typename T, bool MAX> __device__ inline bool beats(T a, T b) { return MAX ? a > b : a < b; }

template <typename T, bool MAX> __device__ inline ExtSum<T> ext_combine(const ExtSum<T> &a, const ExtSum<T> &b) {
    ExtSum<T> r;
    r.has_nan = a.has_nan | b.has_nan;
    r.nanv = a.has_nan ? (b.simd ? canonical_nan(T(0)) : a.nanv) : b.nanv;
    r.simd = a.simd | b.simd;
    if (!b.has_m || (a.has_m && !(beats<T, MAX>(b.m, a.m) || (b.m == a.m && b.flag)))) {
        r.m = a.m;
        r.flag = a.flag;
        r.has_m = a.has_m;
    } else {
        r.m = b.m;
        r.flag = b.flag;
        r.has_m = b.has_m;
    }
    return r;
}
template <typename T, bool MAX> __device__ inline T ext_apply(const ExtSum<T> &s, T p) {
    if (p != p) return s.simd ? canonical_nan(T(0)) : p;
    if (s.has_nan) return s.nanv;
    if (s.has_m && (beats<T, MAX>(s.m, p) || (s.m == p && s.flag))) return s.m;
    return p;
}

// In-order rank of each SIMD lane in the AVX-512 horizontal reduce of
// NumPy 2.2.6's maximum/minimum loop (later rank wins ties).
__constant__ int kLaneRank64[8] = {3, 7, 1, 5, 2, 6, 0, 4};
__constant__ int kLaneRank32[16] = {3, 7, 11, 15, 1, 5, 9, 13, 2, 6, 10, 14, 0, 4, 8, 12};

// Per-thread candidate within a window: best non-NaN (value, key) with
// later key winning ties, first NaN by key, NaN seen by the vector part.
template <typename T> struct ExtCand {
    T m, nanv;
    int64_t km, kn;
    uint32_t has_m, has_nan, vec_nan;
};

template <typename T, bool MAX> __device__ inline void cand_merge(ExtCand<T> &a, const ExtCand<T> &b) {
    if (b.has_m && (!a.has_m || beats<T, MAX>(b.m, a.m) || (b.m == a.m && b.km > a.km))) {
        a.m = b.m;
        a.km = b.km;
        a.has_m = 1;
    }
    if (b.has_nan && (!a.has_nan || b.kn < a.kn)) {
        a.nanv = b.nanv;
        a.kn = b.kn;
        a.has_nan = 1;
    }
    a.vec_nan |= b.vec_nan;
}

// One call of the x86 maximum/minimum loop on `len` elements: the SIMD
// loop (V lanes) when `simd`, else the strided loop (8 scalar accumulators);
// the len % V (or % 8) tail is folded in order after the lanes.
template <typename T> struct ExtGeom {
    int64_t len, nv, per_lane;
    bool simd;
};
template <typename T> __device__ inline ExtGeom<T> ext_geom(int64_t len, bool simd) {
    constexpr int V = sizeof(T) == 8 ? 8 : 16;
    ExtGeom<T> g;
    g.len = len;
    g.simd = simd;
    g.nv = simd ? len - len % V : (len >= 8 ? len - len % 8 : 0);
    g.per_lane = simd ? g.nv / V : g.nv / 8;
    return g;
}
// element e of the call enters the candidate with its position in the
// loop's evaluation order as key
template <typename T, bool MAX> __device__ inline void ext_feed(ExtCand<T> &c, const ExtGeom<T> &g, int64_t e, T v) {
    constexpr int V = sizeof(T) == 8 ? 8 : 16;
    int64_t key;
    if (e < g.nv) {
        const int lane = g.simd ? (int)(e % V) : (int)(e & 7);
        const int rank = g.simd ? (V == 8 ? kLaneRank64[lane] : kLaneRank32[lane]) : lane;
        key = (int64_t)rank * g.per_lane + e / (g.simd ? V : 8);
    } else {
        key = e;
    }
    if (v != v) {
        if (!c.has_nan || key < c.kn) {
            c.has_nan = 1;
            c.kn = key;
            c.nanv = v;
        }
        if (g.simd && e < g.nv) c.vec_nan = 1;
    } else if (!c.has_m || beats<T, MAX>(v, c.m) || (v == c.m && key > c.km)) {
        c.has_m = 1;
        c.m = v;
        c.km = key;
    }
}
template <typename T> __device__ inline void ext_cand_init(ExtCand<T> &c) {
    c.has_m = c.has_nan = c.vec_nan = 0;
    c.km = c.kn = 0;
    c.m = c.nanv = T(0);
}
// the call's summary from the merged candidate of all its elements
template <typename T> __device__ inline ExtSum<T> ext_finish(const ExtCand<T> &b, const ExtGeom<T> &g) {
    constexpr int V = sizeof(T) == 8 ? 8 : 16;
    ExtSum<T> S;
    S.simd = g.simd && g.len >= 1;
    S.has_m = b.has_m;
    S.m = b.m;
    S.flag = g.simd ? (b.km >= (int64_t)(V - 1) * g.per_lane) : 1;
    S.has_nan = b.has_nan;
    S.nanv = (g.simd && b.vec_nan) ? canonical_nan(T(0)) : b.nanv;
    return S;
}
// the empty summary (no element): the identity of ext_combine
template <typename T> __device__ inline ExtSum<T> ext_empty() {
    ExtSum<T> S;
    S.m = S.nanv = T(0);
    S.has_m = S.flag = S.has_nan = S.simd = 0;
    return S;
}

// Summary of window g (block-cooperative; `red` is LDS scratch): every
// element gets its position in the x86 loop's evaluation order, so ties
// between -0 and +0 and NaN payloads come out as NumPy's.
template <typename T, bool MAX> __device__ ExtSum<T> ext_window_summary(const RedArgs &A, int64_t g, ExtCand<T> *red) {
    constexpr int LT = FDT<T>::v;
    const int64_t per_out = A.rsize * A.nwin;
    const int64_t o = g / per_out, rem = g - o * per_out;
    const int64_t r = rem / A.nwin, w = rem - r * A.nwin;
    int64_t ib, ob;
    out_offsets(A, (uint32_t)o, &ib, &ob);
    ib += r_offset(A, r);
    const int64_t s0 = w * A.bufsize;
    const int64_t L = min(A.bufsize, A.wsize - s0);
    // a window inside W's first coalesced run is read in place with that
    // run's stride; one spanning runs goes through NumPy's contiguous buffer
    const bool one_run = (s0 / A.w_shape[0]) == ((s0 + L - 1) / A.w_shape[0]);
    const bool simd = one_run ? (A.w_st[0] == 1) : true;
    const int64_t skip = (!A.has_init && r == 0 && w == 0) ? 1 : 0;  // first element seeds the fold
    const int64_t len = L - skip;
    const ExtGeom<T> geo = ext_geom<T>(len, simd);
    ExtCand<T> c;
    ext_cand_init(c);
    for (int64_t e = threadIdx.x; e < len; e += blockDim.x)
        ext_feed<T, MAX>(c, geo, e, ld<LT>(A.in_dt, A.in, ib + w_offset(A, s0 + skip + e)));
    red[threadIdx.x] = c;
    __syncthreads();
    for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
        if ((int)threadIdx.x < s2) {
            ExtCand<T> a = red[threadIdx.x];
            cand_merge<T, MAX>(a, red[threadIdx.x + s2]);
            red[threadIdx.x] = a;
        }
        __syncthreads();
    }
    const ExtCand<T> b = red[0];
    __syncthreads();
    return ext_finish(b, geo);
}

template <typename T, bool MAX>
__global__ __launch_bounds__(256) void ext_windows(const RedArgs A, int64_t nwindows) {
    __shared__ ExtCand<T> red[256];
    for (int64_t g = blockIdx.x; g < nwindows; g += gridDim.x) {
        const ExtSum<T> S = ext_window_summary<T, MAX>(A, g, red);
        if (threadIdx.x == 0) ((ExtSum<T> *)A.scratch)[g] = S;
    }
}

// Contiguous windows (W one run of stride 1, no outer axes), `wpb`
// consecutive windows of one output per workgroup, summarised in one pass:
// with every window running NumPy's SIMD loop, the ordered combine of their
// summaries reduces to
//   best value   - the best non-NaN element of the run (equal non-zero
//                  values have equal bits, so tie order is invisible);
//   NaN outcome  - canonical NaN if a NaN occurs before the run's last
//                  window or in the vector part of the last window, else the
//                  first NaN of the last window's scalar tail.
// Only a run whose best value is a zero (a -0/+0 tie picks a sign by
// position) is recomputed window by window with ext_window_summary.  Each
// thread keeps its candidate across the run; the workgroup combines once.
template <typename T, bool MAX>
__global__ __launch_bounds__(256) void ext_fast(const RedArgs A, int64_t wpb, int64_t nblk) {
    constexpr int V = sizeof(T) == 8 ? 8 : 16;  // NumPy's AVX-512 lanes
    constexpr int VEC = 16 / sizeof(T);
    constexpr int NV = 8192 / (256 * VEC);      // vectors per thread in a full window
    using VT = T __attribute__((ext_vector_type(VEC)));
    __shared__ ExtCand<T> red[256];
    __shared__ T wm[4], wtv[4];
    __shared__ uint32_t wfl[4];
    __shared__ int64_t wtk[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t gb = blockIdx.x; gb < A.nout * nblk; gb += gridDim.x) {
        const int64_t o = gb / nblk, bk = gb - o * nblk;
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t w0 = bk * wpb, w1 = min(A.nwin, (bk + 1) * wpb);
        T m = T(0), tv = T(0);
        uint32_t has_m = 0, nan_early = 0, vnan_last = 0, any_simd = 0;
        int64_t tk = INT64_MAX;  // first NaN in the last window's tail
        // windows [w0, wf) are full and 16-byte aligned: each is scanned
        // while the next one's loads are in flight
        const T *xb = (const T *)A.in + ib;
        const int64_t wf = (A.bufsize == 8192 && ((uintptr_t)xb % 16) == 0) ? min(w1, A.wsize / 8192) : w0;
        VT v[NV], pf[NV];
        auto load = [&](VT *r, int64_t w) {
#pragma unroll
            for (int k = 0; k < NV; k++)
                r[k] = __builtin_nontemporal_load((const VT *)(xb + w * 8192 + ((int64_t)k * 256 + threadIdx.x) * VEC));
        };
        if (w0 < wf) load(v, w0);
        for (int64_t w = w0; w < w1; w++) {
            const bool last = w == w1 - 1;
            const int64_t s0 = w * A.bufsize;
            const int64_t L = min(A.bufsize, A.wsize - s0);
            const int64_t skip = (!A.has_init && w == 0) ? 1 : 0;  // without initial= element 0 is the seed
            const int64_t len = L - skip, nv = len - len % V;
            any_simd |= len >= 1;
            const T *x = (const T *)A.in + ib + s0;
            auto visit = [&](T v, int64_t e) {  // e: index within the window
                if (e < skip) return;
                const int64_t rel = e - skip;
                if (v != v) {
                    if (!last) nan_early = 1;
                    else if (rel < nv) vnan_last = 1;
                    else if (rel < tk) {
                        tk = rel;
                        tv = v;
                    }
                } else if (!has_m || beats<T, MAX>(v, m)) {
                    m = v;
                    has_m = 1;
                }
            };
            if (w < wf) {
                if (w + 1 < wf) load(pf, w + 1);
                // Usual case, no NaN among this thread's elements: only the
                // best value matters, and the hardware max/min give it (equal
                // non-zero values have equal bits; a zero best is redone
                // exactly below).  Otherwise, or when the thread holds the
                // seed element, every element is visited.
                bool isn = false;
                T mx = v[0][0];
#pragma unroll
                for (int k = 0; k < NV; k++)
#pragma unroll
                    for (int c = 0; c < VEC; c++) {
                        isn |= v[k][c] != v[k][c];
                        mx = MAX ? __builtin_fmax(mx, v[k][c]) : __builtin_fmin(mx, v[k][c]);
                    }
                if (!isn && !(skip && threadIdx.x == 0)) {
                    if (!has_m || beats<T, MAX>(mx, m)) {
                        m = mx;
                        has_m = 1;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < NV; k++)
#pragma unroll
                        for (int c = 0; c < VEC; c++) visit(v[k][c], ((int64_t)k * 256 + threadIdx.x) * VEC + c);
                }
#pragma unroll
                for (int k = 0; k < NV; k++) v[k] = pf[k];
            } else {
                for (int64_t e = threadIdx.x; e < L; e += blockDim.x) visit(x[e], e);
            }
        }
        // workgroup combine: wavefront shuffles, then the 4 wavefronts
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const T om = __shfl_xor(m, off, 64);
            const uint32_t oh = __shfl_xor(has_m, off, 64);
            if (oh && (!has_m || beats<T, MAX>(om, m))) {
                m = om;
                has_m = 1;
            }
            const int64_t otk = __shfl_xor(tk, off, 64);
            const T otv = __shfl_xor(tv, off, 64);
            if (otk < tk) {
                tk = otk;
                tv = otv;
            }
        }
        const uint32_t fl = (__ballot(nan_early) ? 2u : 0u) | (__ballot(vnan_last) ? 4u : 0u) | has_m;
        if (lane == 0) {
            wm[wave] = m;
            wfl[wave] = fl;
            wtk[wave] = tk;
            wtv[wave] = tv;
        }
        __syncthreads();
        T M = T(0), TV = T(0);
        uint32_t FL = 0;
        int64_t TK = INT64_MAX;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
            if ((wfl[q] & 1) && (!(FL & 1) || beats<T, MAX>(wm[q], M))) M = wm[q];
            FL |= wfl[q];
            if (wtk[q] < TK) {
                TK = wtk[q];
                TV = wtv[q];
            }
        }
        __syncthreads();
        ExtSum<T> acc;
        if ((FL & 1) && M == T(0)) {  // a zero best: signs by position, window by window
            acc.has_m = acc.has_nan = acc.simd = acc.flag = 0;
            acc.m = acc.nanv = T(0);
            for (int64_t w = w0; w < w1; w++) {
                const ExtSum<T> S = ext_window_summary<T, MAX>(A, o * A.nwin + w, red);
                acc = ext_combine<T, MAX>(acc, S);
            }
        } else {
            acc.simd = any_simd;
            acc.has_m = FL & 1;
            acc.m = M;
            acc.flag = 1;
            acc.has_nan = (FL & 6) || TK != INT64_MAX;
            acc.nanv = (FL & 6) ? canonical_nan(T(0)) : TV;
        }
        if (threadIdx.x == 0) ((ExtSum<T> *)A.scratch)[gb] = acc;
    }
}

// Per output: ordered combine of its window summaries (block-wide tree that
// keeps left-to-right order), applied to the seed element.
template <typename T, bool MAX> __global__ __launch_bounds__(256) void ext_fold(const RedArgs A, int64_t per_out) {
    constexpr int LT = FDT<T>::v;
    __shared__ ExtSum<T> red[256];
    for (int64_t o = blockIdx.x; o < A.nout; o += gridDim.x) {
        const ExtSum<T> *S = (const ExtSum<T> *)A.scratch + o * per_out;
        const int64_t per_t = (per_out + blockDim.x - 1) / blockDim.x;
        const int64_t b = threadIdx.x * per_t, e = min(per_out, b + per_t);
        ExtSum<T> acc;
        acc.has_m = acc.has_nan = acc.simd = acc.flag = 0;
        acc.m = acc.nanv = T(0);
        for (int64_t i = b; i < e; i++) acc = ext_combine<T, MAX>(acc, S[i]);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = 1; s < (int)blockDim.x; s <<= 1) {
            if ((threadIdx.x % (2 * s)) == 0 && threadIdx.x + s < blockDim.x)
                red[threadIdx.x] = ext_combine<T, MAX>(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            int64_t ib, ob;
            out_offsets(A, (uint32_t)o, &ib, &ob);
            const T seed = A.has_init ? fseed<T>(A) : ld<LT>(A.in_dt, A.in, ib);
            ((T *)A.out)[ob] = ext_apply<T, MAX>(red[0], seed);
        }
        __syncthreads();
    }
}

// W empty: acc = first element; acc = maxn(acc, x) in iterator order.
template <typename T, bool MAX> __global__ __launch_bounds__(256) void ext_narrow(const RedArgs A) {
    constexpr int LT = FDT<T>::v;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        T acc = A.has_init ? fseed<T>(A) : ld<LT>(A.in_dt, A.in, ib);
        for (int64_t r = A.has_init ? 0 : 1; r < A.rsize; r++) {
            const T v = ld<LT>(A.in_dt, A.in, ib + r_offset(A, r));
            acc = MAX ? fmaxn(acc, v) : fminn(acc, v);
        }
        ((T *)A.out)[ob] = acc;
    }
}

// ---------------------------------------------------------------------------
// where= masked reductions on floats (reduce_axis_masked, ufunc.rs:7172, as
// NumPy 2.2.6 evaluates it): the fold starts from initial= (add: else 0;
// maximum/minimum always have one), and inside every bufsize window the
// reduce loop is called once per maximal run of True mask values:
//   add              acc = acc + pairwise(run)
//   maximum/minimum  the window's SIMD (or strided) loop model on the run
// W empty: acc = op(acc, x) for each selected x in iterator order.
// Multiply is the product chain with masked-out factors 1 (prod_*).
// ---------------------------------------------------------------------------
template <typename T> __device__ inline T init_value(const RedArgs &A, T dflt) {
    if (!A.has_init) return dflt;
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, A.init_bits);
    else return __builtin_bit_cast(T, (uint32_t)A.init_bits);
}

template <typename T, int OP> __global__ __launch_bounds__(256) void mask_narrow(const RedArgs A) {
    constexpr int LT = FDT<T>::v;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t mb = mask_k_offset(A, (uint32_t)o);
        T acc = init_value<T>(A, T(0));
        for (int64_t r = 0; r < A.rsize; r++) {
            if (!A.mask[mb + mask_r_offset(A, r)]) continue;
            const T v = ld<LT>(A.in_dt, A.in, ib + r_offset(A, r));
            if constexpr (OP == IONP_ADD) acc = fadd(acc, v);
            else if constexpr (OP == IONP_MAXIMUM) acc = fmaxn(acc, v);
            else acc = fminn(acc, v);
        }
        ((T *)A.out)[ob] = acc;
    }
}

// exclusive prefix sum of v over the workgroup (blockDim.x a multiple of 64);
// *total gets the sum.  wtmp: one int per wavefront (LDS).
__device__ inline int block_excl_scan(int v, int *wtmp, int *total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wtmp[wv] = x;
    __syncthreads();
    int off = 0, tot = 0;
    for (int k = 0; k < nwv; k++) {
        const int t = wtmp[k];
        if (k < wv) off += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return off + x - v;
}

// W not empty: a workgroup per output walks its windows in order.  A
// window's values and mask are staged in LDS; each thread owns the runs
// that start in its chunk of at most 32 positions, so every run but a
// thread's last is short, and a run longer than one pairwise leaf (> 128)
// is always a thread's last.  Short runs are evaluated by their owner;
// long runs by the whole workgroup (pairwise_block / a keyed candidate
// reduction).  Add folds the run sums into acc in run order; max/min
// combine the run summaries in order (ext_combine) and apply the window's
// summary to acc.
template <typename T, int OP> __global__ __launch_bounds__(256) void mask_rows(const RedArgs A) {
    constexpr int LT = FDT<T>::v;
    constexpr bool ADD = OP == IONP_ADD;
    constexpr bool MAX = OP == IONP_MAXIMUM;
    constexpr int NT = 256;
    __shared__ T vals[kPwMaxWin];
    __shared__ uint8_t flags[kPwMaxWin];
    __shared__ int ffalse[NT];  // first masked-out position in each thread's chunk, or -1
    __shared__ int wtmp[NT / 64];
    __shared__ int nlong, l_lo[64], l_n[64], l_key[64];
    __shared__ T runres[ADD ? kPwMaxWin / 2 : 1];  // add: the runs' pairwise sums by ordinal
    __shared__ T racc[ADD ? 1024 : 1], pval[ADD ? 1024 : 1];
    __shared__ int leaf_lo[ADD ? 128 : 1], leaf_n[ADD ? 128 : 1], lcnt[1];
    __shared__ ExtCand<T> red[ADD ? 1 : NT];
    __shared__ ExtSum<T> tsum[ADD ? 1 : NT];  // max/min: long-run summary, then thread totals
    const int tid = threadIdx.x;
    for (int64_t o = blockIdx.x; o < A.nout; o += gridDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t mb = mask_k_offset(A, (uint32_t)o);
        T acc = init_value<T>(A, T(0));  // thread 0 keeps the running value
        for (int64_t r = 0; r < A.rsize; r++) {
            const int64_t rb = ib + r_offset(A, r), mrb = mb + mask_r_offset(A, r);
            for (int64_t w = 0; w < A.nwin; w++) {
                const int64_t s0 = w * A.bufsize;
                const int L = (int)min(A.bufsize, A.wsize - s0);
                for (int e = tid; e < L; e += NT) {
                    const uint8_t f = A.mask[mrb + mask_w_offset(A, s0 + e)] != 0;
                    flags[e] = f;
                    if (f) vals[e] = ld<LT>(A.in_dt, A.in, rb + w_offset(A, s0 + e));
                }
                if (tid == 0) nlong = 0;
                __syncthreads();
                const int C = (L + NT - 1) / NT;
                const int c0 = min(L, tid * C), c1 = min(L, c0 + C);
                int cnt = 0, ff = -1;
                for (int e = c0; e < c1; e++) {
                    cnt += flags[e] && (e == 0 || !flags[e - 1]);
                    if (ff < 0 && !flags[e]) ff = e;
                }
                ffalse[tid] = ff;
                int total;
                int k = block_excl_scan(cnt, wtmp, &total);  // run ordinal of this thread's first run
                // the max/min loop kind of this window (as for unmasked windows)
                const bool one_run = (s0 / A.w_shape[0]) == ((s0 + L - 1) / A.w_shape[0]);
                const bool simd = one_run ? (A.w_st[0] == 1) : true;
                ExtSum<T> pre = ext_empty<T>();
                bool haslong = false;
                for (int e = c0; e < c1; e++) {
                    if (!(flags[e] && (e == 0 || !flags[e - 1]))) continue;
                    int f = e + 1;
                    while (f < c1 && flags[f]) f++;
                    if (f == c1) {  // runs on past the chunk: the next masked-out position
                        f = L;
                        for (int u = tid + 1; u < NT; u++)
                            if (ffalse[u] >= 0) {
                                f = ffalse[u];
                                break;
                            }
                    }
                    const int n = f - e;
                    if (n > 128) {
                        const int i = atomicAdd(&nlong, 1);
                        l_lo[i] = e;
                        l_n[i] = n;
                        l_key[i] = ADD ? k : tid;
                        haslong = true;
                    } else if constexpr (ADD) {
                        runres[k] = pw_leaf<T>([&](int i) { return vals[e + i]; }, n);
                    } else {
                        const ExtGeom<T> g = ext_geom<T>(n, simd);
                        ExtCand<T> c;
                        ext_cand_init(c);
                        for (int i = 0; i < n; i++) ext_feed<T, MAX>(c, g, i, vals[e + i]);
                        pre = ext_combine<T, MAX>(pre, ext_finish(c, g));
                    }
                    k++;
                }
                __syncthreads();
                const int NL = nlong;
                for (int i = 0; i < NL; i++) {
                    const int lo = l_lo[i], n = l_n[i];
                    if constexpr (ADD) {
                        const T sres = pairwise_block<T>([&](int j) { return vals[lo + j]; }, n, racc, pval, leaf_lo,
                                                         leaf_n, lcnt);
                        if (tid == 0) runres[l_key[i]] = sres;
                    } else {
                        const ExtGeom<T> g = ext_geom<T>(n, simd);
                        ExtCand<T> c;
                        ext_cand_init(c);
                        for (int j = tid; j < n; j += NT) ext_feed<T, MAX>(c, g, j, vals[lo + j]);
                        red[tid] = c;
                        __syncthreads();
                        for (int s2 = NT / 2; s2 > 0; s2 >>= 1) {
                            if (tid < s2) {
                                ExtCand<T> a = red[tid];
                                cand_merge<T, MAX>(a, red[tid + s2]);
                                red[tid] = a;
                            }
                            __syncthreads();
                        }
                        if (tid == 0) tsum[l_key[i]] = ext_finish(red[0], g);
                    }
                    __syncthreads();
                }
                if constexpr (ADD) {
                    if (tid == 0)
                        for (int i = 0; i < total; i++) acc = fadd(acc, runres[i]);
                } else {
                    const ExtSum<T> mine = haslong ? ext_combine<T, MAX>(pre, tsum[tid]) : pre;
                    __syncthreads();
                    tsum[tid] = mine;
                    __syncthreads();
                    for (int s2 = 1; s2 < NT; s2 <<= 1) {  // ordered: left operand first
                        if ((tid & (2 * s2 - 1)) == 0) tsum[tid] = ext_combine<T, MAX>(tsum[tid], tsum[tid + s2]);
                        __syncthreads();
                    }
                    if (tid == 0) acc = ext_apply<T, MAX>(tsum[0], acc);
                }
                __syncthreads();
            }
        }
        if (tid == 0) ((T *)A.out)[ob] = acc;
    }
}

// ---------------------------------------------------------------------------
// order-free reductions: integer / bool add, multiply, maximum, minimum
// ---------------------------------------------------------------------------
template <typename AT, int OP> __device__ inline AT assoc(AT a, AT b) {
    if constexpr (OP == IONP_ADD) return (AT)((uint64_t)a + (uint64_t)b);  // two's-complement wrap
    else if constexpr (OP == IONP_MULTIPLY) return (AT)((uint64_t)a * (uint64_t)b);
    else if constexpr (OP == IONP_MAXIMUM) return a > b ? a : b;
    else return a < b ? a : b;
}
// the op's identity on an integer accumulator (the starting value of a
// where= masked fold, whose elements may all be masked out)
template <typename AT, int OP> __device__ constexpr AT int_identity() {
    if constexpr (OP == IONP_ADD) return AT(0);
    else if constexpr (OP == IONP_MULTIPLY) return AT(1);
    else if constexpr (OP == IONP_MAXIMUM) return std::numeric_limits<AT>::lowest();
    else return std::numeric_limits<AT>::max();
}
// a group's result combined with initial= (order-free: once, at the end)
template <typename AT, int OP> __device__ inline AT with_init(const RedArgs &A, AT acc) {
    return A.has_init ? assoc<AT, OP>(acc, (AT)A.init_bits) : acc;
}

template <typename AT, int LT> __device__ inline AT load_acc(int dt, const char *base, int64_t off) {
    return (AT)ld<LT>(dt, base, off);
}

// AT: int64_t / uint64_t accumulator; LT: dtype the elements are cast to;
// MASK: where= masked (elements whose mask byte is 0 are skipped).
template <typename AT, int LT, int OP, bool MASK = false>
__global__ __launch_bounds__(256) void int_reduce(const RedArgs A, int64_t seg, int64_t nseg) {
    __shared__ AT red[256];
    using ST = typename dt_type<LT>::T;
    for (int64_t g = blockIdx.x; g < A.nout * nseg; g += gridDim.x) {
        const int64_t o = g / nseg, sg = g - o * nseg;
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t mb = MASK ? mask_k_offset(A, (uint32_t)o) : 0;
        const int64_t q0 = sg * seg, q1 = min(A.group, q0 + seg);
        AT acc = MASK ? int_identity<AT, OP>()
                 : OP == IONP_ADD ? AT(0) : OP == IONP_MULTIPLY ? AT(1) : (AT)load_acc<AT, LT>(A.in_dt, A.in, ib + group_offset(A, q0));
        for (int64_t q = q0 + threadIdx.x; q < q1; q += blockDim.x) {
            if constexpr (MASK)
                if (!A.mask[mb + mask_group_offset(A, q)]) continue;
            acc = assoc<AT, OP>(acc, (AT)load_acc<ST, LT>(A.in_dt, A.in, ib + group_offset(A, q)));
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = assoc<AT, OP>(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (nseg == 1) ((ST *)A.out)[ob] = int_result<LT>(with_init<AT, OP>(A, red[0]));
            else ((AT *)A.scratch)[g] = red[0];
        }
        __syncthreads();
    }
}
// per output, a workgroup combines its segment partials (order-free)
template <typename AT, int LT, int OP> __global__ __launch_bounds__(256) void int_reduce_fold(const RedArgs A, int64_t nseg) {
    using ST = typename dt_type<LT>::T;
    __shared__ AT red[256];
    for (int64_t o = blockIdx.x; o < A.nout; o += gridDim.x) {
        const AT *p = (const AT *)A.scratch + o * nseg;
        AT acc = OP == IONP_ADD ? AT(0) : OP == IONP_MULTIPLY ? AT(1) : p[0];
        for (int64_t i = threadIdx.x; i < nseg; i += blockDim.x) acc = assoc<AT, OP>(acc, p[i]);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
            if ((int)threadIdx.x < s2) red[threadIdx.x] = assoc<AT, OP>(red[threadIdx.x], red[threadIdx.x + s2]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            int64_t ib, ob;
            out_offsets(A, (uint32_t)o, &ib, &ob);
            ((ST *)A.out)[ob] = int_result<LT>(with_init<AT, OP>(A, red[0]));
        }
        __syncthreads();
    }
}

// contiguous groups (one run of stride 1 per output) in the input dtype:
// 16-byte loads, values widened to the accumulator
template <typename AT, typename IT, typename ST, int OP>
__global__ __launch_bounds__(256) void int_reduce_contig(const RedArgs A, int64_t seg, int64_t nseg) {
    constexpr int VEC = 16 / sizeof(IT);
    using V = IT __attribute__((ext_vector_type(VEC)));
    __shared__ AT red[256];
    for (int64_t g = blockIdx.x; g < A.nout * nseg; g += gridDim.x) {
        const int64_t o = g / nseg, sg = g - o * nseg;
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const IT *x = (const IT *)A.in + ib;
        const int64_t q0 = sg * seg, q1 = min(A.group, q0 + seg);
        AT acc = OP == IONP_ADD ? AT(0) : OP == IONP_MULTIPLY ? AT(1) : (AT)x[q0];
        // aligned vector body
        int64_t v0 = q0, v1 = q1;
        while (v0 < q1 && ((uintptr_t)(x + v0) % 16)) v0++;
        v1 = v0 + (q1 - v0) / VEC * VEC;
        // four vectors per thread in flight, then the rest one at a time
        constexpr int U = 4;
        const int64_t step = (int64_t)blockDim.x * VEC;
        int64_t q = v0 + (int64_t)threadIdx.x * VEC;
        for (; q + (U - 1) * step < v1; q += U * step) {
            V v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load((const V *)(x + q + u * step));
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int k = 0; k < VEC; k++) acc = assoc<AT, OP>(acc, (AT)v[u][k]);
        }
        for (; q < v1; q += step) {
            const V v = __builtin_nontemporal_load((const V *)(x + q));
#pragma unroll
            for (int k = 0; k < VEC; k++) acc = assoc<AT, OP>(acc, (AT)v[k]);
        }
        for (int64_t q = q0 + threadIdx.x; q < v0; q += blockDim.x) acc = assoc<AT, OP>(acc, (AT)x[q]);
        for (int64_t q = v1 + threadIdx.x; q < q1; q += blockDim.x) acc = assoc<AT, OP>(acc, (AT)x[q]);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
            if ((int)threadIdx.x < s2) red[threadIdx.x] = assoc<AT, OP>(red[threadIdx.x], red[threadIdx.x + s2]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (nseg == 1) ((ST *)A.out)[ob] = (ST)with_init<AT, OP>(A, red[0]);  // never a bool output (see run_int)
            else ((AT *)A.scratch)[g] = red[0];
        }
        __syncthreads();
    }
}

// narrow layouts: one thread per output, elements in any order
template <typename AT, int LT, int OP, bool MASK = false>
__global__ __launch_bounds__(256) void int_reduce_cols(const RedArgs A) {
    using ST = typename dt_type<LT>::T;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t mb = MASK ? mask_k_offset(A, (uint32_t)o) : 0;
        AT acc = MASK ? int_identity<AT, OP>()
                 : OP == IONP_ADD ? AT(0) : OP == IONP_MULTIPLY ? AT(1) : (AT)load_acc<ST, LT>(A.in_dt, A.in, ib);
        for (int64_t q = 0; q < A.group; q++) {
            if constexpr (MASK)
                if (!A.mask[mb + mask_group_offset(A, q)]) continue;
            acc = assoc<AT, OP>(acc, (AT)load_acc<ST, LT>(A.in_dt, A.in, ib + group_offset(A, q)));
        }
        ((ST *)A.out)[ob] = int_result<LT>(with_init<AT, OP>(A, acc));
    }
}

// zero-size groups: initial= if given (multiply's identity 1 is passed the
// same way), else add's identity 0
template <typename ST> __global__ void fill_zero(const RedArgs A) {
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        ((ST *)A.out)[ob] = A.has_init ? (ST)A.init_bits : ST(0);  // initial= or add's identity
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int grid_cap(int64_t items, int bpc) {
    int64_t g = (int64_t)bpc * num_cus();
    if (g > items) g = items;
    return (int)(g > 0 ? g : 1);
}

template <typename T, typename IT> static void launch_narrow_t(const RedArgs &A, int bpc, hipStream_t s) {
    if (narrow_wave_ok(A))
        hipLaunchKernelGGL((pw_narrow_wave<T, IT>), dim3(grid_cap((A.nout + 3) / 4, bpc)), dim3(256), 0, s, A);
    else
        hipLaunchKernelGGL((pw_narrow<T, IT>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
}
template <typename T> static void launch_narrow(const RedArgs &A, int bpc, hipStream_t s) {
    switch (A.in_dt) {
        case IONP_BOOL: launch_narrow_t<T, bool>(A, bpc, s); break;
        case IONP_INT8: launch_narrow_t<T, int8_t>(A, bpc, s); break;
        case IONP_UINT8: launch_narrow_t<T, uint8_t>(A, bpc, s); break;
        case IONP_INT16: launch_narrow_t<T, int16_t>(A, bpc, s); break;
        case IONP_UINT16: launch_narrow_t<T, uint16_t>(A, bpc, s); break;
        case IONP_INT32: launch_narrow_t<T, int32_t>(A, bpc, s); break;
        case IONP_UINT32: launch_narrow_t<T, uint32_t>(A, bpc, s); break;
        case IONP_INT64: launch_narrow_t<T, int64_t>(A, bpc, s); break;
        case IONP_UINT64: launch_narrow_t<T, uint64_t>(A, bpc, s); break;
        case IONP_FLOAT32: launch_narrow_t<T, float>(A, bpc, s); break;
        default: launch_narrow_t<T, double>(A, bpc, s); break;
    }
}

template <typename T, typename IT> static void launch_prod_t(const RedArgs &A, int bpc, hipStream_t s) {
    const bool wave = A.nout <= 8192 && A.group >= 256;
    const dim3 gw(grid_cap((A.nout + 3) / 4, bpc)), gc(grid_cap((A.nout + 255) / 256, bpc)), blk(256);
    if (A.mask) {
        if (wave) hipLaunchKernelGGL((prod_wave<T, IT, true>), gw, blk, 0, s, A);
        else hipLaunchKernelGGL((prod_cols<T, IT, true>), gc, blk, 0, s, A);
    } else {
        if (wave) hipLaunchKernelGGL((prod_wave<T, IT, false>), gw, blk, 0, s, A);
        else hipLaunchKernelGGL((prod_cols<T, IT, false>), gc, blk, 0, s, A);
    }
}
template <typename T> static void launch_prod(const RedArgs &A, int bpc, hipStream_t s) {
    switch (A.in_dt) {
        case IONP_BOOL: launch_prod_t<T, bool>(A, bpc, s); break;
        case IONP_INT8: launch_prod_t<T, int8_t>(A, bpc, s); break;
        case IONP_UINT8: launch_prod_t<T, uint8_t>(A, bpc, s); break;
        case IONP_INT16: launch_prod_t<T, int16_t>(A, bpc, s); break;
        case IONP_UINT16: launch_prod_t<T, uint16_t>(A, bpc, s); break;
        case IONP_INT32: launch_prod_t<T, int32_t>(A, bpc, s); break;
        case IONP_UINT32: launch_prod_t<T, uint32_t>(A, bpc, s); break;
        case IONP_INT64: launch_prod_t<T, int64_t>(A, bpc, s); break;
        case IONP_UINT64: launch_prod_t<T, uint64_t>(A, bpc, s); break;
        case IONP_FLOAT32: launch_prod_t<T, float>(A, bpc, s); break;
        default: launch_prod_t<T, double>(A, bpc, s); break;
    }
}

static ionp_status launched() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

template <typename AT, int LT, int OP, bool MASK>
static ionp_status run_int(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (int_cols_ok(A)) {
        hipLaunchKernelGGL((int_reduce_cols<AT, LT, OP, MASK>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0,
                           s, A);
        return launched();
    }
    // segments of at least 65536 elements, and no more of them than one
    // round of workgroups (fewer partials for int_reduce_fold)
    const int64_t slots = (int64_t)num_cus() * bpc;
    const int64_t seg = std::max<int64_t>(65536, ((A.group + slots - 1) / slots + 4095) / 4096 * 4096);
    const int64_t nseg = (A.group + seg - 1) / seg;
    if (nseg > 1) {
        A.scratch = scratch_get((size_t)(A.nout * nseg) * sizeof(AT), s);
        if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for reduction scratch");
    }
    using ST = typename dt_type<LT>::T;
    const dim3 grid(grid_cap(A.nout * nseg, bpc)), blk(256);
    // raw input values only where that equals casting each to the output
    // dtype first: integer add and multiply (wrap modulo the output width
    // either way) or no cast; bool outputs (logical or / and) take int_reduce
    if (!MASK && A.nw == 1 && A.w_st[0] == 1 && A.nr == 0 && A.in_dt != IONP_BOOL && A.out_dt != IONP_BOOL &&
        !is_real_float(A.in_dt) &&  // float input: each element goes through the C cast (ld)
        A.group < (1ll << 40) && (OP == IONP_ADD || OP == IONP_MULTIPLY || A.in_dt == A.out_dt)) {
        switch (A.in_dt) {  // one contiguous run per output: vector loads
            case IONP_INT8: hipLaunchKernelGGL((int_reduce_contig<AT, int8_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
            case IONP_UINT8: hipLaunchKernelGGL((int_reduce_contig<AT, uint8_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
            case IONP_INT16: hipLaunchKernelGGL((int_reduce_contig<AT, int16_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
            case IONP_UINT16: hipLaunchKernelGGL((int_reduce_contig<AT, uint16_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
            case IONP_INT32: hipLaunchKernelGGL((int_reduce_contig<AT, int32_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
            case IONP_UINT32: hipLaunchKernelGGL((int_reduce_contig<AT, uint32_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
            case IONP_INT64: hipLaunchKernelGGL((int_reduce_contig<AT, int64_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
            default: hipLaunchKernelGGL((int_reduce_contig<AT, uint64_t, ST, OP>), grid, blk, 0, s, A, seg, nseg); break;
        }
    } else {
        hipLaunchKernelGGL((int_reduce<AT, LT, OP, MASK>), grid, blk, 0, s, A, seg, nseg);
    }
    if (nseg > 1) hipLaunchKernelGGL((int_reduce_fold<AT, LT, OP>), dim3(grid_cap(A.nout, bpc)), dim3(256), 0, s, A, nseg);
    return launched();
}

template <int LT, int OP, bool MASK> static ionp_status run_int_t(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    if constexpr (LT == IONP_UINT64 || LT == IONP_UINT32 || LT == IONP_UINT16 || LT == IONP_UINT8 || LT == IONP_BOOL)
        return run_int<uint64_t, LT, OP, MASK>(A, cfg, s);
    else
        return run_int<int64_t, LT, OP, MASK>(A, cfg, s);
}

template <int OP, bool MASK = false> static ionp_status run_int_dispatch(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    switch (A.out_dt) {
        case IONP_BOOL: return run_int_t<IONP_BOOL, OP, MASK>(A, cfg, s);
        case IONP_INT8: return run_int_t<IONP_INT8, OP, MASK>(A, cfg, s);
        case IONP_UINT8: return run_int_t<IONP_UINT8, OP, MASK>(A, cfg, s);
        case IONP_INT16: return run_int_t<IONP_INT16, OP, MASK>(A, cfg, s);
        case IONP_UINT16: return run_int_t<IONP_UINT16, OP, MASK>(A, cfg, s);
        case IONP_INT32: return run_int_t<IONP_INT32, OP, MASK>(A, cfg, s);
        case IONP_UINT32: return run_int_t<IONP_UINT32, OP, MASK>(A, cfg, s);
        case IONP_INT64: return run_int_t<IONP_INT64, OP, MASK>(A, cfg, s);
        case IONP_UINT64: return run_int_t<IONP_UINT64, OP, MASK>(A, cfg, s);
        default: return set_error(IONP_ERR_NO_LOOP, "no integer reduce loop");
    }
}

// where= masked reductions (the mask is set in A)
template <typename T> static ionp_status run_masked_float(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (A.op == IONP_MULTIPLY) {
        launch_prod<T>(A, bpc, s);
        return launched();
    }
    if (A.nw == 0) {
        const dim3 g(grid_cap((A.nout + 255) / 256, bpc)), b(256);
        if (A.op == IONP_ADD) hipLaunchKernelGGL((mask_narrow<T, IONP_ADD>), g, b, 0, s, A);
        else if (A.op == IONP_MAXIMUM) hipLaunchKernelGGL((mask_narrow<T, IONP_MAXIMUM>), g, b, 0, s, A);
        else hipLaunchKernelGGL((mask_narrow<T, IONP_MINIMUM>), g, b, 0, s, A);
        return launched();
    }
    if (A.bufsize > kPwMaxWin) return set_error(IONP_ERR_UNSUPPORTED, "bufsize above 8192");
    const dim3 g(grid_cap(A.nout, bpc)), b(256);
    if (A.op == IONP_ADD) hipLaunchKernelGGL((mask_rows<T, IONP_ADD>), g, b, 0, s, A);
    else if (A.op == IONP_MAXIMUM) hipLaunchKernelGGL((mask_rows<T, IONP_MAXIMUM>), g, b, 0, s, A);
    else hipLaunchKernelGGL((mask_rows<T, IONP_MINIMUM>), g, b, 0, s, A);
    return launched();
}
static ionp_status run_masked(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    if (is_float_dt(A.out_dt))
        return A.out_dt == IONP_FLOAT64 ? run_masked_float<double>(A, cfg, s) : run_masked_float<float>(A, cfg, s);
    switch (A.op) {
        case IONP_ADD: return run_int_dispatch<IONP_ADD, true>(A, cfg, s);
        case IONP_MULTIPLY: return run_int_dispatch<IONP_MULTIPLY, true>(A, cfg, s);
        case IONP_MAXIMUM: return run_int_dispatch<IONP_MAXIMUM, true>(A, cfg, s);
        default: return run_int_dispatch<IONP_MINIMUM, true>(A, cfg, s);
    }
}

// Launch of pw_stream (one workgroup per CU: the dynamic LDS request below
// puts a workgroup above half of a CU's 160 KB).  So a folder has its CU to
// itself (sharing one with producers, its chain fell behind them: 2^28
// float64 elements in 358 against 337 us, measured on MI355X), and every
// producer is resident at once, which the folders' waits rely on.  Producers
// on 7/8 of the CUs stream faster than on all of them (float32
// (4096,65536) rows: 165 against 170 us); the other CUs host the folders.
// The launch shape is fixed by residency, so ionp_config does not apply.
constexpr size_t kPsPadLds = 16u << 10;
// The folders' give-up record of each device (see PwStreamState): `dirty`
// in device memory, `err` in fine-grained host memory.  Allocated with the
// library's fold slots; an installed ScratchArena refills its slots on
// every use and reports through `err` only if it already exists.
struct FoldStatus {
    uint32_t *dirty = nullptr, *err = nullptr;
};
static FoldStatus g_fold[64];
static std::atomic<int64_t> g_fold_stall{-1};
static std::atomic<uint32_t> g_fold_polls{1u << 26};

static FoldStatus fold_status(bool alloc) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return FoldStatus();
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    FoldStatus &F = g_fold[dev];
    if (alloc && !F.dirty) {
        void *d = nullptr, *h = nullptr;
        if (hipMalloc(&d, 256) != hipSuccess) return FoldStatus();
        if (hipHostMalloc(&h, 64, hipHostMallocCoherent) != hipSuccess || hipMemset(d, 0, 256) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            (void)hipFree(d);
            if (h) (void)hipHostFree(h);
            return FoldStatus();
        }
        *(volatile uint32_t *)h = 0;
        F.dirty = (uint32_t *)d;
        F.err = (uint32_t *)h;
    }
    return F;
}

bool fold_error_take() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
    uint32_t *e = g_fold[dev].err;
    return e && __atomic_exchange_n(e, 0u, __ATOMIC_ACQ_REL) != 0;
}

void fold_debug_stall(int64_t window, uint32_t polls) {
    g_fold_stall = window;
    g_fold_polls = polls;
}

template <typename T> static ionp_status run_pw_stream(RedArgs &A, hipStream_t s) {
    PwStreamState F;
    F.nfull = A.wsize / 8192;
    const int64_t has_tail = A.wsize % 8192 != 0;
    F.nwin = F.nfull + has_tail;
    F.folders = (A.nout <= 64 && F.nwin > 8) ? (int32_t)A.nout : 0;
    F.poll_limit = g_fold_polls.load();
    F.stall = g_fold_stall.load();
    F.dirty = nullptr;
    F.err = nullptr;
    const int64_t nw_total = A.nout * F.nwin;
    // folders poll the slots for the sentinel; pw_fold runs after the launch
    // and needs none
    T *part = F.folders ? slots_get<T>(nw_total, s) : (T *)scratch_get((size_t)nw_total * sizeof(T), s);
    if (!part) return set_error(IONP_ERR_DEVICE, "out of device memory for reduction scratch");
    if (F.folders) {
        const FoldStatus fs = fold_status(!t_arena);
        F.err = fs.err;
        if (!t_arena) {
            if (!fs.dirty) return set_error(IONP_ERR_DEVICE, "out of memory for the fold status words");
            F.dirty = fs.dirty;
            // a folder that gave up in an earlier call may have left late
            // values in the slots: refill them all first (a no-op otherwise)
            hipLaunchKernelGGL(slots_refill<T>, dim3(1), dim3(256), 0, s, part, (int64_t)(g_slots_bytes / sizeof(T)),
                               F.dirty);
        }
    }
    F.partials = part;
    A.scratch = part;
    if (has_tail)  // the last, partial window of each output (generic tree)
        hipLaunchKernelGGL(pw_windows<T>, dim3(grid_cap(A.nout, 8)), dim3(kPwThreads), 0, s, A, A.nout, 0, F.nwin);
    const int64_t nwindows = A.nout * F.nfull;
    static int resident = 0;  // workgroups per CU (1 with the pad)
    if (!resident) {
        int b = 0;
        resident = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, pw_stream<T, true>, 64 * kPsWaves, kPsPadLds) == hipSuccess && b > 0
                       ? b
                       : 1;
    }
    const int64_t cus = num_cus();
    int64_t prod = std::min<int64_t>((nwindows + kPsWaves - 1) / kPsWaves, resident * (cus * 7 / 8));
    prod = std::max<int64_t>(1, std::min<int64_t>(prod, resident * cus - F.folders));
    const dim3 grid((unsigned)(F.folders + prod)), blk(64 * kPsWaves);
    if ((double)A.nout * (double)A.wsize * sizeof(T) > (double)kInfinityCacheBytes)  // streamed once: nontemporal
        hipLaunchKernelGGL((pw_stream<T, true>), grid, blk, kPsPadLds, s, A, F, nwindows);
    else
        hipLaunchKernelGGL((pw_stream<T, false>), grid, blk, kPsPadLds, s, A, F, nwindows);
    if (!F.folders) hipLaunchKernelGGL(pw_fold<T>, dim3(grid_cap((A.nout + 255) / 256, 8)), dim3(256), 0, s, A);
    const ionp_status st = launched();
    if (st != IONP_OK && F.folders) slots_dirty();  // slots may hold values: refill before the next use
    return st;
}

template <typename T> static ionp_status run_float_add(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (A.nw == 0) {
        launch_narrow<T>(A, bpc, s);
        return launched();
    }
    // contiguous rows with at least one full 8192-element window, native dtype;
    // either few windows per output or few outputs (one folder each)
    const bool rows_ok = A.nw == 1 && A.w_st[0] == 1 && A.rsize == 1 && A.bufsize == 8192 && A.wsize >= 8192 &&
                         A.in_dt == A.out_dt && ((uintptr_t)A.in % sizeof(T)) == 0 &&
                         A.wsize / 8192 * A.nout < (1ll << 31);
    if (rows_ok)
        return run_pw_stream<T>(A, s);
    if (pw_small_ok(A)) {
        hipLaunchKernelGGL(pw_small<T>, dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
        return launched();
    }
    const int64_t nwindows = A.nout * A.rsize * A.nwin;
    const int direct = A.rsize == 1 && A.nwin == 1;
    if (!direct) {
        A.scratch = scratch_get((size_t)nwindows * sizeof(T), s);
        if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for reduction scratch");
    }
    hipLaunchKernelGGL(pw_windows<T>, dim3(grid_cap(nwindows, bpc)), dim3(kPwThreads), 0, s, A, nwindows, direct, (int64_t)0);
    if (!direct) hipLaunchKernelGGL(pw_fold<T>, dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
    return launched();
}

template <typename T, bool MAX> static ionp_status run_float_ext(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (A.nw == 0) {
        hipLaunchKernelGGL((ext_narrow<T, MAX>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
        return launched();
    }
    const int64_t nwindows = A.nout * A.rsize * A.nwin;
    A.scratch = scratch_get((size_t)nwindows * sizeof(ExtSum<T>), s);
    if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for reduction scratch");
    if (A.nw == 1 && A.w_st[0] == 1 && A.nr == 0 && A.in_dt == A.out_dt) {
        // one round of resident workgroups, each folding a run of windows
        // (fewer, longer runs: fewer summaries for ext_fold)
        static int resident = 0;
        if (!resident) {
            int b = 0;
            resident = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, ext_fast<T, MAX>, 256, 0) == hipSuccess && b > 0 ? b : 2;
        }
        const int64_t want = (int64_t)num_cus() * std::min(resident, bpc);
        int64_t wpb = (nwindows + want - 1) / want;
        if (wpb < 1) wpb = 1;
        if (wpb > A.nwin) wpb = A.nwin;
        const int64_t nblk = (A.nwin + wpb - 1) / wpb;
        hipLaunchKernelGGL((ext_fast<T, MAX>), dim3(grid_cap(A.nout * nblk, bpc)), dim3(256), 0, s, A, wpb, nblk);
        hipLaunchKernelGGL((ext_fold<T, MAX>), dim3(grid_cap(A.nout, bpc)), dim3(256), 0, s, A, nblk);
    } else {
        hipLaunchKernelGGL((ext_windows<T, MAX>), dim3(grid_cap(nwindows, bpc)), dim3(256), 0, s, A, nwindows);
        hipLaunchKernelGGL((ext_fold<T, MAX>), dim3(grid_cap(A.nout, bpc)), dim3(256), 0, s, A, A.rsize * A.nwin);
    }
    return launched();
}

// ---------------------------------------------------------------------------
// float16 outputs (NumPy's HALF loops compute in float32): add -- per
// bufsize window acc = half(float(acc) + pairwise_f32(window)), the reduce
// loop storing its float accumulator back to the float16 output after
// every call; multiply -- per window a float product chain from float(acc),
// rounded at the window's end; maximum / minimum -- element by element on
// float16 values (hstep).  W empty: each element is its own call, rounded
// every step.  A workgroup per output; the serial parts run in thread 0.
// ---------------------------------------------------------------------------
// MASK (where=): NumPy calls the reduce loop once per maximal run of True
// mask values inside each window (W empty: per selected element); each call
// keeps its float32 accumulator and stores it back to float16 at its end,
// as without a mask (reduce_axis_masked_f16_widen, ufunc.rs:7414, widens
// the same way).  Thread 0 finds the runs of a window in order.
template <int OP, bool MASK> __global__ __launch_bounds__(256) void half_reduce(const RedArgs A) {
    __shared__ float racc[1024], pval[1024];
    __shared__ int leaf_lo[128], leaf_n[128], lcnt[1];
    __shared__ int run_lo, run_hi;
    const int tid = threadIdx.x;
    const int isz = itemsize(A.in_dt);
    auto xh = [&](int64_t off) { return to_half_bits(A.in_dt, A.in + off * isz); };
    for (int64_t o = blockIdx.x; o < A.nout; o += gridDim.x) {
        int64_t ib, ob;
        out_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t mb = MASK ? mask_k_offset(A, (uint32_t)o) : 0;
        uint16_t acc = A.has_init ? (uint16_t)A.init_bits : (OP == IONP_MULTIPLY ? (uint16_t)0x3c00 : (uint16_t)0);
        bool first = !A.has_init && (OP == IONP_MAXIMUM || OP == IONP_MINIMUM);  // first element seeds
        for (int64_t r = 0; r < A.rsize; r++) {
            const int64_t rb = ib + r_offset(A, r);
            const int64_t rmb = MASK ? mb + mask_r_offset(A, r) : 0;
            if (A.nw == 0) {
                if (tid == 0 && (!MASK || A.mask[rmb])) {
                    const uint16_t x = xh(rb);
                    acc = first ? x : hstep(OP, acc, x);
                    first = false;
                }
                continue;
            }
            for (int64_t w = 0; w < A.nwin; w++) {
                const int64_t s0 = w * A.bufsize;
                const int L = (int)min(A.bufsize, A.wsize - s0);
                // the calls of this window: [lo, hi) runs (the whole window unmasked)
                int lo = 0;
                for (;;) {
                    int hi = L;
                    if constexpr (MASK) {
                        auto sel = [&](int i) { return A.mask[rmb + mask_w_offset(A, s0 + i)] != 0; };
                        if (tid == 0) {
                            int a = lo;
                            while (a < L && !sel(a)) a++;
                            int b = a;
                            while (b < L && sel(b)) b++;
                            run_lo = a;
                            run_hi = b;
                        }
                        __syncthreads();
                        lo = run_lo;
                        hi = run_hi;
                        __syncthreads();  // read by every thread before thread 0 writes the next run
                        if (lo >= L) break;
                    }
                    const int base = lo;
                    if constexpr (OP == IONP_ADD) {
                        const float pw = pairwise_block<float>([&](int i) { return h2f(xh(rb + w_offset(A, s0 + base + i))); },
                                                               hi - lo, racc, pval, leaf_lo, leaf_n, lcnt);
                        if (tid == 0) {
                            uint32_t fl = 0;
                            acc = f2h(fadd(h2f(acc), pw), &fl);
                        }
                        __syncthreads();
                    } else if (tid == 0) {
                        if constexpr (OP == IONP_MULTIPLY) {
                            float f = h2f(acc);
                            for (int i = lo; i < hi; i++) f = fmul(f, h2f(xh(rb + w_offset(A, s0 + i))));
                            uint32_t fl = 0;
                            acc = f2h(f, &fl);
                        } else {
                            for (int i = lo; i < hi; i++) {
                                const uint16_t x = xh(rb + w_offset(A, s0 + i));
                                acc = first ? x : hstep(OP, acc, x);
                                first = false;
                            }
                        }
                    }
                    if (!MASK) break;
                    lo = hi;
                }
            }
        }
        if (tid == 0) ((uint16_t *)A.out)[ob] = acc;
        __syncthreads();
    }
}

template <bool MASK> static void launch_half(const RedArgs &A, dim3 g, dim3 b, hipStream_t s) {
    switch (A.op) {
        case IONP_ADD: hipLaunchKernelGGL((half_reduce<IONP_ADD, MASK>), g, b, 0, s, A); break;
        case IONP_MULTIPLY: hipLaunchKernelGGL((half_reduce<IONP_MULTIPLY, MASK>), g, b, 0, s, A); break;
        case IONP_MAXIMUM: hipLaunchKernelGGL((half_reduce<IONP_MAXIMUM, MASK>), g, b, 0, s, A); break;
        default: hipLaunchKernelGGL((half_reduce<IONP_MINIMUM, MASK>), g, b, 0, s, A); break;
    }
}

static ionp_status run_half_reduce(RedArgs &A, const ionp_config *cfg, hipStream_t s) {
    if (A.bufsize > kPwMaxWin) return set_error(IONP_ERR_UNSUPPORTED, "bufsize above 8192");
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    const dim3 g(grid_cap(A.nout, bpc)), b(256);
    if (A.mask) launch_half<true>(A, g, b, s);
    else launch_half<false>(A, g, b, s);
    return launched();
}

uint64_t init_bits_of(int out_dt, const void *p) {
    switch (out_dt) {
        case IONP_BOOL: return *(const uint8_t *)p != 0;
        case IONP_INT8: return (uint64_t)(int64_t)*(const int8_t *)p;
        case IONP_UINT8: return *(const uint8_t *)p;
        case IONP_INT16: return (uint64_t)(int64_t)*(const int16_t *)p;
        case IONP_UINT16:
        case IONP_FLOAT16: return *(const uint16_t *)p;
        case IONP_INT32: return (uint64_t)(int64_t)*(const int32_t *)p;
        case IONP_UINT32: return *(const uint32_t *)p;
        case IONP_FLOAT32: return *(const uint32_t *)p;
        default: return *(const uint64_t *)p;  // int64, uint64, float64 bits
    }
}

ionp_status run_reduce(const ReducePlan &p, int64_t bufsize, const void *initial, const ionp_config *cfg,
                       hipStream_t s) {
    if (p.nout == 0) return IONP_OK;
    RedArgs A;
    std::memset(&A, 0, sizeof A);
    A.in_dt = p.in_dt;
    A.out_dt = p.out_dt;
    A.op = p.op;
    A.nk = p.nk;
    A.nw = p.nw;
    A.nr = p.nr;
    A.nout = p.nout;
    A.wsize = p.wsize;
    A.rsize = p.rsize;
    A.group = p.group;
    A.bufsize = bufsize;
    A.nwin = p.wsize > 0 ? (p.wsize + bufsize - 1) / bufsize : 0;
    A.in = p.in;
    A.out = p.out;
    A.has_init = initial != nullptr;
    A.init_bits = initial ? init_bits_of(p.out_dt, initial) : 0;
    if (p.nout >= (1ll << 31) || p.wsize >= (1ll << 31) * 8 || p.rsize >= (1ll << 31))
        return set_error(IONP_ERR_UNSUPPORTED, "reduction extent too large");
    A.mask = p.mask;
    for (int d = 0; d < p.nk; d++) {
        A.k_shape[d] = p.k_shape[d];
        A.k_in_st[d] = p.k_in_st[d];
        A.k_out_st[d] = p.k_out_st[d];
        A.k_m_st[d] = p.k_m_st[d];
        A.k_div[d] = FastDiv((uint32_t)p.k_shape[d]);
    }
    for (int d = 0; d < p.nw; d++) {
        A.w_shape[d] = p.w_shape[d];
        A.w_st[d] = p.w_st[d];
        A.w_m_st[d] = p.w_m_st[d];
        A.w_div[d] = FastDiv((uint32_t)(p.w_shape[d] < (1ll << 31) ? p.w_shape[d] : 1));
    }
    if (p.nw > 1 && p.wsize >= (1ll << 31)) return set_error(IONP_ERR_UNSUPPORTED, "reduction extent too large");
    for (int d = 0; d < p.nr; d++) {
        A.r_shape[d] = p.r_shape[d];
        A.r_st[d] = p.r_st[d];
        A.r_m_st[d] = p.r_m_st[d];
        A.r_div[d] = FastDiv((uint32_t)p.r_shape[d]);
    }
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (p.group == 0) {
        if (p.op != IONP_ADD && p.op != IONP_MULTIPLY && !initial)
            return set_error(IONP_ERR_VALUE, std::string("zero-size array to reduction operation ") +
                                                               op_name(p.op) + " which has no identity");
        if (p.op == IONP_MULTIPLY && !initial) {  // multiply's identity 1 in the output dtype
            A.has_init = 1;
            A.init_bits = p.out_dt == IONP_FLOAT64   ? 0x3ff0000000000000ull
                          : p.out_dt == IONP_FLOAT32 ? 0x3f800000ull
                          : p.out_dt == IONP_FLOAT16 ? 0x3c00ull
                                                     : 1ull;
        }
        const int g = grid_cap((p.nout + 255) / 256, bpc);
        switch (itemsize(p.out_dt)) {
            case 1: hipLaunchKernelGGL(fill_zero<uint8_t>, dim3(g), dim3(256), 0, s, A); break;
            case 2: hipLaunchKernelGGL(fill_zero<uint16_t>, dim3(g), dim3(256), 0, s, A); break;
            case 4: hipLaunchKernelGGL(fill_zero<uint32_t>, dim3(g), dim3(256), 0, s, A); break;
            default: hipLaunchKernelGGL(fill_zero<uint64_t>, dim3(g), dim3(256), 0, s, A); break;
        }
        return launched();
    }
    if (p.out_dt == IONP_FLOAT16) return run_half_reduce(A, cfg, s);
    if (A.mask) return run_masked(A, cfg, s);
    if (is_float_dt(p.out_dt)) {
        if (p.op == IONP_ADD) {
            if (bufsize > kPwMaxWin) return set_error(IONP_ERR_UNSUPPORTED, "bufsize above 8192");
            return p.out_dt == IONP_FLOAT64 ? run_float_add<double>(A, cfg, s) : run_float_add<float>(A, cfg, s);
        }
        if (p.op == IONP_MULTIPLY) {
            if (p.out_dt == IONP_FLOAT64) launch_prod<double>(A, bpc, s);
            else launch_prod<float>(A, bpc, s);
            return launched();
        }
        if (p.op == IONP_MAXIMUM)
            return p.out_dt == IONP_FLOAT64 ? run_float_ext<double, true>(A, cfg, s) : run_float_ext<float, true>(A, cfg, s);
        return p.out_dt == IONP_FLOAT64 ? run_float_ext<double, false>(A, cfg, s) : run_float_ext<float, false>(A, cfg, s);
    }
    switch (p.op) {
        case IONP_ADD: return run_int_dispatch<IONP_ADD>(A, cfg, s);
        case IONP_MULTIPLY: return run_int_dispatch<IONP_MULTIPLY>(A, cfg, s);
        case IONP_MAXIMUM: return run_int_dispatch<IONP_MAXIMUM>(A, cfg, s);
        default: return run_int_dispatch<IONP_MINIMUM>(A, cfg, s);
    }
}

// ---------------------------------------------------------------------------
// argmax / argmin (reduce_axis_argext, ufunc.rs:8037; first NaN wins, then
// the first strict extreme)
// ---------------------------------------------------------------------------
template <typename T, bool MAX> struct ArgBest {
    T v;
    int64_t i;
    uint32_t has, nan;
};
template <typename T, bool MAX> __device__ inline void arg_merge(ArgBest<T, MAX> &a, const ArgBest<T, MAX> &b) {
    if (!b.has) return;
    if (!a.has) { a = b; return; }
    if (a.nan || b.nan) {
        if (b.nan && (!a.nan || b.i < a.i)) a = b;
        return;
    }
    if (beats<T, MAX>(b.v, a.v) || (b.v == a.v && b.i < a.i)) a = b;
}
template <typename T> __device__ inline bool is_nan_any(T v) {
    if constexpr (std::is_floating_point<T>::value) return v != v;
    else return false;
}

__device__ inline void arg_offsets(const ArgArgs &A, uint32_t o, int64_t *ib, int64_t *ob) {
    int64_t a = 0, b = 0;
    for (int d = 0; d < A.nk; d++) {
        uint32_t c;
        if (d + 1 < A.nk) {
            const uint32_t q = A.k_div[d].div(o);
            c = o - q * (uint32_t)A.k_shape[d];
            o = q;
        } else {
            c = o;
        }
        a += (int64_t)c * A.k_in_st[d];
        b += (int64_t)c * A.k_out_st[d];
    }
    *ib = a;
    *ob = b;
}
__device__ inline int64_t arg_l_offset(const ArgArgs &A, int64_t i) {
    if (A.nl == 1) return i * A.l_st[0];
    int64_t off = 0;
    uint32_t idx = (uint32_t)i;
    for (int d = 0; d < A.nl; d++) {
        uint32_t c;
        if (d + 1 < A.nl) {
            const uint32_t q = A.l_div[d].div(idx);
            c = idx - q * (uint32_t)A.l_shape[d];
            idx = q;
        } else {
            c = idx;
        }
        off += (int64_t)c * A.l_st[d];
    }
    return off;
}

template <typename T, bool MAX> __global__ __launch_bounds__(256) void arg_rows(const ArgArgs A) {
    __shared__ ArgBest<T, MAX> red[256];
    const T *x = (const T *)A.in;
    for (int64_t g = blockIdx.x; g < A.nout * A.nseg; g += gridDim.x) {
        const int64_t o = g / A.nseg, sg = g - o * A.nseg;
        int64_t ib, ob;
        arg_offsets(A, (uint32_t)o, &ib, &ob);
        const int64_t i0 = sg * A.seg, i1 = min(A.len, i0 + A.seg);
        ArgBest<T, MAX> b;
        b.has = 0;
        b.nan = 0;
        b.i = 0;
        b.v = T(0);
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            const T v = x[ib + arg_l_offset(A, i)];
            ArgBest<T, MAX> c;
            c.v = v;
            c.i = i;
            c.has = 1;
            c.nan = is_nan_any(v);
            arg_merge(b, c);
        }
        red[threadIdx.x] = b;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                ArgBest<T, MAX> a = red[threadIdx.x];
                arg_merge(a, red[threadIdx.x + s]);
                red[threadIdx.x] = a;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (A.nseg == 1) ((int64_t *)A.out)[ob] = red[0].i;
            else ((ArgBest<T, MAX> *)A.scratch)[g] = red[0];
        }
        __syncthreads();
    }
}
// per output, a workgroup merges its segment candidates (order-free merge)
template <typename T, bool MAX> __global__ __launch_bounds__(256) void arg_fold(const ArgArgs A) {
    __shared__ ArgBest<T, MAX> red[256];
    for (int64_t o = blockIdx.x; o < A.nout; o += gridDim.x) {
        const ArgBest<T, MAX> *p = (const ArgBest<T, MAX> *)A.scratch + o * A.nseg;
        ArgBest<T, MAX> b;
        b.has = 0;
        b.nan = 0;
        b.i = 0;
        b.v = T(0);
        for (int64_t k = threadIdx.x; k < A.nseg; k += blockDim.x) arg_merge(b, p[k]);
        red[threadIdx.x] = b;
        __syncthreads();
        for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
            if ((int)threadIdx.x < s2) {
                ArgBest<T, MAX> c = red[threadIdx.x];
                arg_merge(c, red[threadIdx.x + s2]);
                red[threadIdx.x] = c;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            int64_t ib, ob;
            arg_offsets(A, (uint32_t)o, &ib, &ob);
            ((int64_t *)A.out)[ob] = red[0].i;
        }
        __syncthreads();
    }
}

// contiguous scan (stride 1): 16-byte loads; each thread visits its indices
// in increasing order, so a strict compare keeps the first extreme and the
// first NaN ends its scan
template <typename T, bool MAX> __global__ __launch_bounds__(256) void arg_contig(const ArgArgs A) {
    constexpr int VEC = 16 / sizeof(T) > 0 ? 16 / sizeof(T) : 1;
    using V = T __attribute__((ext_vector_type(VEC)));
    __shared__ ArgBest<T, MAX> red[256];
    for (int64_t g = blockIdx.x; g < A.nout * A.nseg; g += gridDim.x) {
        const int64_t o = g / A.nseg, sg = g - o * A.nseg;
        int64_t ib, ob;
        arg_offsets(A, (uint32_t)o, &ib, &ob);
        const T *x = (const T *)A.in + ib;
        const int64_t i0 = sg * A.seg, i1 = min(A.len, i0 + A.seg);
        int64_t v0 = i0;
        while (v0 < i1 && ((uintptr_t)(x + v0) % 16)) v0++;
        const int64_t v1 = v0 + (i1 - v0) / VEC * VEC;
        ArgBest<T, MAX> b;
        b.has = 0;
        b.nan = 0;
        b.i = 0;
        b.v = T(0);
        auto visit = [&](T v, int64_t i) {
            if (b.nan) return;
            if (is_nan_any(v)) {
                b.v = v;
                b.i = i;
                b.nan = 1;
                b.has = 1;
            } else if (!b.has || beats<T, MAX>(v, b.v)) {
                b.v = v;
                b.i = i;
                b.has = 1;
            }
        };
        for (int64_t i = i0 + threadIdx.x; i < v0; i += blockDim.x) visit(x[i], i);
        for (int64_t i = v0 + (int64_t)threadIdx.x * VEC; i < v1; i += (int64_t)blockDim.x * VEC) {
            const V v = __builtin_nontemporal_load((const V *)(x + i));
#pragma unroll
            for (int k = 0; k < VEC; k++) visit(v[k], i + k);
        }
        for (int64_t i = v1 + threadIdx.x; i < i1; i += blockDim.x) visit(x[i], i);
        red[threadIdx.x] = b;
        __syncthreads();
        for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
            if ((int)threadIdx.x < s2) {
                ArgBest<T, MAX> c = red[threadIdx.x];
                arg_merge(c, red[threadIdx.x + s2]);
                red[threadIdx.x] = c;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (A.nseg == 1) ((int64_t *)A.out)[ob] = red[0].i;
            else ((ArgBest<T, MAX> *)A.scratch)[g] = red[0];
        }
        __syncthreads();
    }
}

template <typename T, bool MAX> __global__ __launch_bounds__(256) void arg_cols(const ArgArgs A) {
    const T *x = (const T *)A.in;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < A.nout; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t ib, ob;
        arg_offsets(A, (uint32_t)o, &ib, &ob);
        T best = x[ib];
        int64_t bi = 0;
        if (!is_nan_any(best)) {
            for (int64_t i = 1; i < A.len; i++) {
                const T v = x[ib + i * A.l_st[0]];
                if (is_nan_any(v)) { bi = i; break; }
                if (beats<T, MAX>(v, best)) { best = v; bi = i; }
            }
        }
        ((int64_t *)A.out)[ob] = bi;
    }
}

template <typename T, bool MAX> static ionp_status run_arg_t(ArgArgs &A, const ionp_config *cfg, hipStream_t s) {
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    const bool cols = A.nl == 1 && A.l_st[0] != 1 && A.nout >= 4096 && A.len <= 65536;
    if (cols) {
        hipLaunchKernelGGL((arg_cols<T, MAX>), dim3(grid_cap((A.nout + 255) / 256, bpc)), dim3(256), 0, s, A);
        return launched();
    }
    // segments of at least 65536 indices, and no more of them than one
    // round of workgroups (fewer partials for arg_fold)
    const int64_t slots = (int64_t)num_cus() * bpc;
    A.seg = std::max<int64_t>(65536, ((A.len + slots - 1) / slots + 4095) / 4096 * 4096);
    A.nseg = (A.len + A.seg - 1) / A.seg;
    if (A.nseg > 1) {
        A.scratch = scratch_get((size_t)(A.nout * A.nseg) * sizeof(ArgBest<T, MAX>), s);
        if (!A.scratch) return set_error(IONP_ERR_DEVICE, "out of device memory for argext scratch");
    }
    if (A.nl == 1 && A.l_st[0] == 1 && sizeof(T) <= 16)
        hipLaunchKernelGGL((arg_contig<T, MAX>), dim3(grid_cap(A.nout * A.nseg, bpc)), dim3(256), 0, s, A);
    else
        hipLaunchKernelGGL((arg_rows<T, MAX>), dim3(grid_cap(A.nout * A.nseg, bpc)), dim3(256), 0, s, A);
    if (A.nseg > 1) hipLaunchKernelGGL((arg_fold<T, MAX>), dim3(grid_cap(A.nout, bpc)), dim3(256), 0, s, A);
    return launched();
}

template <bool MAX> static ionp_status run_arg_m(ArgArgs &A, const ionp_config *cfg, hipStream_t s) {
    switch (A.in_dt) {
        case IONP_BOOL: return run_arg_t<uint8_t, MAX>(A, cfg, s);
        case IONP_INT8: return run_arg_t<int8_t, MAX>(A, cfg, s);
        case IONP_UINT8: return run_arg_t<uint8_t, MAX>(A, cfg, s);
        case IONP_INT16: return run_arg_t<int16_t, MAX>(A, cfg, s);
        case IONP_UINT16: return run_arg_t<uint16_t, MAX>(A, cfg, s);
        case IONP_INT32: return run_arg_t<int32_t, MAX>(A, cfg, s);
        case IONP_UINT32: return run_arg_t<uint32_t, MAX>(A, cfg, s);
        case IONP_INT64: return run_arg_t<int64_t, MAX>(A, cfg, s);
        case IONP_UINT64: return run_arg_t<uint64_t, MAX>(A, cfg, s);
        case IONP_FLOAT32: return run_arg_t<float, MAX>(A, cfg, s);
        default: return run_arg_t<double, MAX>(A, cfg, s);
    }
}

ionp_status run_argext(ArgArgs &A, const ionp_config *cfg, hipStream_t s) {
    if (A.nout == 0) return IONP_OK;
    if (A.nout >= (1ll << 31) || (A.nl > 1 && A.len >= (1ll << 31)))
        return set_error(IONP_ERR_UNSUPPORTED, "argext extent too large");
    for (int d = 0; d < A.nk; d++) A.k_div[d] = FastDiv((uint32_t)A.k_shape[d]);
    for (int d = 0; d < A.nl; d++) A.l_div[d] = FastDiv((uint32_t)(A.l_shape[d] < (1ll << 31) ? A.l_shape[d] : 1));
    return A.ismax ? run_arg_m<true>(A, cfg, s) : run_arg_m<false>(A, cfg, s);
}

// ---------------------------------------------------------------------------
// ufunc.reduceat (reduceat_binary_axis, ufunc.rs:8298-8460) as NumPy 2.2.6
// evaluates it: output (f, k) is x[i_k] when i_{k+1} <= i_k (the last index
// runs to the axis end), else the reduce loop seeded with x[i_k] and called
// once on the rest [i_k + 1, i_{k+1}) -- no buffer windows:
//   float add       x[i_k] + pairwise(rest), the whole rest one tree
//   float multiply  sequential product (rat_prod)
//   float max/min   the SIMD loop model on a contiguous rest, the strided
//                   loop model otherwise
//   integers        order-free, in the output dtype
// ---------------------------------------------------------------------------
__device__ inline void rat_offsets(const RatArgs &A, uint32_t f, int64_t *ib, int64_t *ob) {
    int64_t a = 0, b = 0;
    for (int d = 0; d < A.nf; d++) {
        uint32_t c;
        if (d + 1 < A.nf) {
            const uint32_t q = A.fdiv[d].div(f);
            c = f - q * (uint32_t)A.fshape[d];
            f = q;
        } else {
            c = f;
        }
        a += (int64_t)c * A.fin[d];
        b += (int64_t)c * A.fout[d];
    }
    *ib = a;
    *ob = b;
}
__device__ inline int64_t pw_split64(int64_t n) { return (n >> 1) & ~(int64_t)7; }

// pairwise(x[0..n)) with the workgroup for any n: nodes of at most 8192
// elements go through pairwise_block in tree order, the nodes above them
// are combined on a stack in LDS (written by thread 0; every thread runs
// the same walk).  Depth <= 32 for n < 2^44.
template <typename T> struct PwStack {
    int64_t lo[32], n[32];
    T left[32];
    int state[32];  // 0: left child pending, 1: left value held
};
template <typename T, typename Get>
__device__ T pairwise_any(const Get &get, int64_t n, T *racc, T *val, int *leaf_lo, int *leaf_n, int *cnt,
                          PwStack<T> &S) {
    const bool t0 = threadIdx.x == 0;
    int sp = 0;
    if (t0) {
        S.lo[0] = 0;
        S.n[0] = n;
        S.state[0] = 0;
    }
    __syncthreads();
    T ret;
    for (;;) {
        const int64_t lo = S.lo[sp], m = S.n[sp];
        if (m <= kPwMaxWin) {
            ret = pairwise_block<T>([&](int i) { return get(lo + i); }, (int)m, racc, val, leaf_lo, leaf_n, cnt);
            __syncthreads();  // every thread has read the result before the scratch is reused
            for (;;) {  // ret: the value of node sp
                if (sp == 0) return ret;
                --sp;
                if (S.state[sp] == 0) {  // left child done: keep it, descend into the right one
                    const int64_t plo = S.lo[sp], pn = S.n[sp];
                    __syncthreads();
                    if (t0) {
                        const int64_t h = pw_split64(pn);
                        S.left[sp] = ret;
                        S.state[sp] = 1;
                        S.lo[sp + 1] = plo + h;
                        S.n[sp + 1] = pn - h;
                        S.state[sp + 1] = 0;
                    }
                    __syncthreads();
                    ++sp;
                    break;
                }
                ret = fadd(S.left[sp], ret);
            }
        } else {
            __syncthreads();
            if (t0) {
                S.lo[sp + 1] = lo;
                S.n[sp + 1] = pw_split64(m);
                S.state[sp + 1] = 0;
            }
            __syncthreads();
            ++sp;
        }
    }
}

// a workgroup per output (f, k); LT: the output dtype
template <int LT, int OP> __global__ __launch_bounds__(256) void rat_block(const RatArgs A) {
    using T = typename dt_type<LT>::T;
    constexpr bool FLT = std::is_floating_point<T>::value;
    constexpr bool PW = FLT && OP == IONP_ADD;
    constexpr bool EXT = FLT && (OP == IONP_MAXIMUM || OP == IONP_MINIMUM);
    using AT = typename std::conditional<FLT, T,
                                         typename std::conditional<std::is_signed<T>::value, int64_t, uint64_t>::type>::type;
    __shared__ T racc[PW ? 1024 : 1], pval[PW ? 1024 : 1];
    __shared__ int leaf_lo[PW ? 128 : 1], leaf_n[PW ? 128 : 1], lcnt[1];
    __shared__ PwStack<T> pstack;
    __shared__ ExtCand<T> red[EXT ? 256 : 1];
    __shared__ AT ired[(!FLT) ? 256 : 1];
    const int tid = threadIdx.x;
    const int64_t nout = A.nfib * A.nidx;
    for (int64_t g = blockIdx.x; g < nout; g += gridDim.x) {
        const int64_t f = g / A.nidx, k = g - f * A.nidx;
        int64_t ib, ob;
        rat_offsets(A, (uint32_t)f, &ib, &ob);
        const int64_t a = A.idx[k], b = k + 1 < A.nidx ? A.idx[k + 1] : A.len;
        const int64_t m = b > a ? b - a - 1 : 0;
        const T seed = ld<LT>(A.in_dt, A.in, ib + a * A.in_st);
        auto get = [&](int64_t i) { return ld<LT>(A.in_dt, A.in, ib + (a + 1 + i) * A.in_st); };
        T res = seed;
        if (m > 0) {
            if constexpr (PW) {
                res = fadd(seed, pairwise_any<T>(get, m, racc, pval, leaf_lo, leaf_n, lcnt, pstack));
            } else if constexpr (EXT) {
                constexpr bool MAX = OP == IONP_MAXIMUM;
                const ExtGeom<T> geo = ext_geom<T>(m, A.in_st == 1);
                ExtCand<T> c;
                ext_cand_init(c);
                for (int64_t e = tid; e < m; e += blockDim.x) ext_feed<T, MAX>(c, geo, e, get(e));
                red[tid] = c;
                __syncthreads();
                for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
                    if (tid < s2) {
                        ExtCand<T> x = red[tid];
                        cand_merge<T, MAX>(x, red[tid + s2]);
                        red[tid] = x;
                    }
                    __syncthreads();
                }
                res = ext_apply<T, MAX>(ext_finish(red[0], geo), seed);
                __syncthreads();
            } else if constexpr (!FLT) {
                AT acc = int_identity<AT, OP>();
                for (int64_t e = tid; e < m; e += blockDim.x) acc = assoc<AT, OP>(acc, (AT)get(e));
                ired[tid] = acc;
                __syncthreads();
                for (int s2 = blockDim.x / 2; s2 > 0; s2 >>= 1) {
                    if (tid < s2) ired[tid] = assoc<AT, OP>(ired[tid], ired[tid + s2]);
                    __syncthreads();
                }
                res = int_result<LT>(assoc<AT, OP>((AT)seed, ired[0]));
                __syncthreads();
            }
        }
        if (tid == 0) ((T *)A.out)[ob + k * A.out_st] = res;
    }
}

// float multiply: a thread per output, the sequential product (plain IEEE
// multiplies, redone with fmul's x86 NaN rules when the result is NaN)
template <typename T> __global__ __launch_bounds__(256) void rat_prod(const RatArgs A) {
    constexpr int LT = FDT<T>::v;
    const int64_t nout = A.nfib * A.nidx;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nout; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = g / A.nidx, k = g - f * A.nidx;
        int64_t ib, ob;
        rat_offsets(A, (uint32_t)f, &ib, &ob);
        const int64_t a = A.idx[k], b = k + 1 < A.nidx ? A.idx[k + 1] : A.len;
        const T seed = ld<LT>(A.in_dt, A.in, ib + a * A.in_st);
        T acc = seed;
        for (int64_t i = a + 1; i < b; i++) acc = acc * ld<LT>(A.in_dt, A.in, ib + i * A.in_st);
        if (acc != acc) {
            acc = seed;
            for (int64_t i = a + 1; i < b; i++) acc = fmul(acc, ld<LT>(A.in_dt, A.in, ib + i * A.in_st));
        }
        ((T *)A.out)[ob + k * A.out_st] = acc;
    }
}

// float16 outputs: x[i] seeds the call, add x[i] + pairwise_f32(rest) and
// multiply a float product chain, each rounded to float16 once at the end
// of the call; maximum / minimum element by element (hstep)
template <int OP> __global__ __launch_bounds__(256) void rat_half(const RatArgs A) {
    __shared__ float racc[1024], pval[1024];
    __shared__ int leaf_lo[128], leaf_n[128], lcnt[1];
    __shared__ PwStack<float> pstack;
    const int tid = threadIdx.x;
    const int isz = itemsize(A.in_dt);
    const int64_t nout = A.nfib * A.nidx;
    for (int64_t g = blockIdx.x; g < nout; g += gridDim.x) {
        const int64_t f = g / A.nidx, k = g - f * A.nidx;
        int64_t ib, ob;
        rat_offsets(A, (uint32_t)f, &ib, &ob);
        const int64_t a = A.idx[k], b = k + 1 < A.nidx ? A.idx[k + 1] : A.len;
        const int64_t m = b > a ? b - a - 1 : 0;
        auto xh = [&](int64_t i) { return to_half_bits(A.in_dt, A.in + (ib + i * A.in_st) * isz); };
        uint16_t res = xh(a);
        if (m > 0) {
            if constexpr (OP == IONP_ADD) {
                const float pw = pairwise_any<float>([&](int64_t i) { return h2f(xh(a + 1 + i)); }, m, racc, pval, leaf_lo,
                                                     leaf_n, lcnt, pstack);
                uint32_t fl = 0;
                res = f2h(fadd(h2f(res), pw), &fl);
            } else if (tid == 0) {
                if constexpr (OP == IONP_MULTIPLY) {
                    float p = h2f(res);
                    for (int64_t i = a + 1; i < b; i++) p = fmul(p, h2f(xh(i)));
                    uint32_t fl = 0;
                    res = f2h(p, &fl);
                } else {
                    for (int64_t i = a + 1; i < b; i++) res = hstep(OP, res, xh(i));
                }
            }
        }
        if (tid == 0) ((uint16_t *)A.out)[ob + k * A.out_st] = res;
        __syncthreads();
    }
}

template <int LT> static void launch_rat(const RatArgs &A, int bpc, hipStream_t s) {
    const int64_t nout = A.nfib * A.nidx;
    const dim3 g(grid_cap(nout, bpc)), b(256);
    switch (A.op) {
        case IONP_ADD: hipLaunchKernelGGL((rat_block<LT, IONP_ADD>), g, b, 0, s, A); break;
        case IONP_MULTIPLY:
            if constexpr (LT == IONP_FLOAT64 || LT == IONP_FLOAT32)
                hipLaunchKernelGGL((rat_prod<typename dt_type<LT>::T>), dim3(grid_cap((nout + 255) / 256, bpc)), b, 0, s,
                                   A);
            else
                hipLaunchKernelGGL((rat_block<LT, IONP_MULTIPLY>), g, b, 0, s, A);
            break;
        case IONP_MAXIMUM: hipLaunchKernelGGL((rat_block<LT, IONP_MAXIMUM>), g, b, 0, s, A); break;
        default: hipLaunchKernelGGL((rat_block<LT, IONP_MINIMUM>), g, b, 0, s, A); break;
    }
}

ionp_status run_reduceat(RatArgs &A, const int64_t *indices, const ionp_config *cfg, hipStream_t s) {
    if (A.nfib == 0 || A.nidx == 0) return IONP_OK;
    if (A.nfib * A.nidx >= (1ll << 40) || A.nfib >= (1ll << 31))
        return set_error(IONP_ERR_UNSUPPORTED, "reduceat extent too large");
    for (int d = 0; d < A.nf; d++) A.fdiv[d] = FastDiv((uint32_t)A.fshape[d]);
    int64_t *didx = (int64_t *)scratch_get((size_t)A.nidx * sizeof(int64_t), s);
    if (!didx) return set_error(IONP_ERR_DEVICE, "out of device memory for reduceat indices");
    if (hipMemcpyAsync(didx, indices, (size_t)A.nidx * sizeof(int64_t), hipMemcpyHostToDevice, s) != hipSuccess)
        return set_error(IONP_ERR_DEVICE, "copying reduceat indices failed");
    A.idx = didx;
    const int bpc = (cfg && cfg->blocks_per_cu > 0) ? cfg->blocks_per_cu : 8;
    if (A.out_dt == IONP_FLOAT16) {
        const dim3 g(grid_cap(A.nfib * A.nidx, bpc)), b(256);
        switch (A.op) {
            case IONP_ADD: hipLaunchKernelGGL(rat_half<IONP_ADD>, g, b, 0, s, A); break;
            case IONP_MULTIPLY: hipLaunchKernelGGL(rat_half<IONP_MULTIPLY>, g, b, 0, s, A); break;
            case IONP_MAXIMUM: hipLaunchKernelGGL(rat_half<IONP_MAXIMUM>, g, b, 0, s, A); break;
            default: hipLaunchKernelGGL(rat_half<IONP_MINIMUM>, g, b, 0, s, A); break;
        }
        return launched();
    }
    switch (A.out_dt) {
        case IONP_BOOL: launch_rat<IONP_BOOL>(A, bpc, s); break;
        case IONP_INT8: launch_rat<IONP_INT8>(A, bpc, s); break;
        case IONP_UINT8: launch_rat<IONP_UINT8>(A, bpc, s); break;
        case IONP_INT16: launch_rat<IONP_INT16>(A, bpc, s); break;
        case IONP_UINT16: launch_rat<IONP_UINT16>(A, bpc, s); break;
        case IONP_INT32: launch_rat<IONP_INT32>(A, bpc, s); break;
        case IONP_UINT32: launch_rat<IONP_UINT32>(A, bpc, s); break;
        case IONP_INT64: launch_rat<IONP_INT64>(A, bpc, s); break;
        case IONP_UINT64: launch_rat<IONP_UINT64>(A, bpc, s); break;
        case IONP_FLOAT32: launch_rat<IONP_FLOAT32>(A, bpc, s); break;
        default: launch_rat<IONP_FLOAT64>(A, bpc, s); break;
    }
    return launched();
}

}  // namespace ionp
