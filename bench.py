"""Benchmark of the MI355X ufunc / reduction backend.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]

A step is one pass of the large broadcast elementwise workload:
    float64 X(8192,8192) * c(8192,1), float32 X * c,
    float32 X.T[::2] * c[::2]  (non-contiguous view, (4096,8192)),
    int64 I(8192,8192) + j(8192,1) (wrapping add),
X ~ normal(0,1), c ~ uniform[0.5,2), I, j ~ uniform[-2^62, 2^62), all bit-exact
with NumPy.  `value` is the step's algorithmic HBM bytes (every operand read
once, the output written once) divided by the step time, inputs resident in
HBM.  One JSON line is printed; `roofline` covers the dominant kernel (the
float64 multiply), timed with HIP events on the stream it runs on, and
`cpu_baseline` times the scalar C oracle on the same four operations, on one
core and (OpenMP) on every host core.  `extra` carries the other elementwise
and scan workloads (float64 broadcast add, the float64 view, A + B.T,
cumsum), the reduction workloads (float64 add.reduce of 2^28 elements,
float32 (4096,65536) sums, maximum/argmax, also with NaN at rate 1e-6), the
uneven batch and the learned configuration.
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM = 8.0e12


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-extra", action="store_true")
    p.add_argument("--no-sweep", action="store_true",
                   help="skip the learned configuration's live sweep (its ~150k timed launches swamp a kernel trace)")
    p.add_argument("--reductions", action="store_true",
                   help="only run the reduction workloads and print their timings (the workload of the "
                        "reduction kernel profiles)")
    p.add_argument("--kernels", action="store_true",
                   help="only run the elementwise and scan workloads beside the step (broadcast add, "
                        "X.T[::2] view, A + B.T, cumsum) and print their timings")
    p.add_argument("--roofline-pmc", action="store_true",
                   help="only run the dominant kernel and a plain copy of known size, 5 times each "
                        "(the workload of the rocprofv3 FETCH_SIZE / WRITE_SIZE passes)")
    return p.parse_args()


def roofline_pmc_workload():
    """ew_rows<double,multiply> on X(8192,8192) * c(8192,1), then torch's
    device copy of 2^26 float64 (536,870,912 B read and written), 5 times
    each after the launch configuration has settled (the last 5 multiply
    launches are the ones measured): the counter passes scale the
    multiply's FETCH_SIZE / WRITE_SIZE by what the same counters read for
    the copy."""
    import torch
    import anionpy_amd as ap
    torch.cuda.set_device(0)
    ap.seterr(all="ignore")
    X = ap.asarray(np.random.default_rng(1).standard_normal((8192, 8192)))
    c = ap.asarray(np.random.default_rng(2).uniform(0.5, 2, (8192, 1)))
    out = ap.multiply(X, c)
    src = torch.ones(2 ** 26, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    # the launch configuration settles as in the timed bench run
    settled = settle_policy(torch, lambda: ap.multiply(X, c, out=out))
    for _ in range(5):
        ap.multiply(X, c, out=out)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"multiply_bytes": X.nbytes + c.nbytes + out.nbytes, "copy_bytes_each_way": src.numel() * 8,
                      "launches": 5, "settle_calls": settled, "config": launch_config(ap, ap.multiply, X, c, out)}),
          flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


@contextlib.contextmanager
def quiet_gc(torch, warm):
    """A timed region without Python garbage-collection pauses: collected
    first, the collector off inside (a pause of tens of ms inside a loop of
    launches leaves the GPU idle and inflates the time).  `warm` runs after
    the collection and before the region, and the device synchronises: the
    first launches after a collection are slow on the host (measured on
    MI355X: 10 float64 8192 x 8192 adds queued in 300-350 us instead of
    200-220 us), which idles the GPU at the start of the region and added
    about 10 us per call to a 10-call loop."""
    gc.collect()
    gc.disable()
    try:
        warm()
        torch.cuda.synchronize()
        yield
    finally:
        gc.enable()


def main():
    args = parse()
    if args.roofline_pmc:
        roofline_pmc_workload()
        return
    if args.kernels:
        import torch
        import anionpy_amd as ap
        torch.cuda.set_device(0)
        ap.seterr(all="ignore")
        res = elementwise_extras(ap, torch)
        # calibration for counter passes: torch's device copy of 2^26
        # float64 (536,870,912 B read and written), 5 times
        src = torch.ones(2 ** 26, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps(res), flush=True)
        return
    if args.reductions:
        import torch
        import anionpy_amd as ap
        torch.cuda.set_device(0)
        ap.seterr(all="ignore")
        print(json.dumps(reduction_extras(ap, torch)), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; multi-GPU sharding is a design "
                                   "only and not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import anionpy_amd as ap

    torch.cuda.set_device(0)
    # IEEE flag tracking (np.errstate, NumPy's default modes) synchronises
    # after every float call; the timed workload runs without it and its
    # cost is reported separately (extra.fpe_tracking)
    ap.seterr(all="ignore")
    log("generating inputs")
    X = np.random.default_rng(1).standard_normal((8192, 8192))
    c = np.random.default_rng(2).uniform(0.5, 2, (8192, 1))
    X64, c64 = ap.asarray(X), ap.asarray(c)
    X32, c32 = ap.asarray(X.astype(np.float32)), ap.asarray(c.astype(np.float32))
    del X
    I = ap.asarray(np.random.default_rng(3).integers(-2 ** 62, 2 ** 62, (8192, 8192)))
    j = ap.asarray(np.random.default_rng(8).integers(-2 ** 62, 2 ** 62, (8192, 1)))
    V32, v32 = X32.T[::2], c32[::2]

    ops = [("mul_f64", ap.multiply, X64, c64), ("mul_f32", ap.multiply, X32, c32),
           ("view_f32", ap.multiply, V32, v32), ("add_i64", ap.add, I, j)]
    outs = [f(a, b) for _, f, a, b in ops]  # preallocated outputs (K-order layouts)

    def alg_bytes(a, b, o):
        return a.nbytes + b.nbytes + o.nbytes

    step_bytes = sum(alg_bytes(a, b, o) for (_, _, a, b), o in zip(ops, outs))

    def step():
        for (_, f, a, b), o in zip(ops, outs):
            f(a, b, out=o)

    # the learned launch configuration settles first: each step layout's
    # proposal is evaluated and verified against the heuristic on live calls
    # (a trial); untimed steps run until no layout is pending or under trial
    settle = settle_policy(torch, step)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    with quiet_gc(torch, step):
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    value = step_bytes * args.steps / dt / 1e9

    # per-op kernel time with HIP events on the launch stream
    per_op = {}
    for (name, f, a, b), o in zip(ops, outs):
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        with quiet_gc(torch, lambda: f(a, b, out=o)):
            e0.record(s)
            for _ in range(reps):
                f(a, b, out=o)
            e1.record(s)
            e1.synchronize()
        t = e0.elapsed_time(e1) / reps * 1e-3
        per_op[name] = {"us": round(t * 1e6, 1), "GB/s": round(alg_bytes(a, b, o) / t / 1e9, 1),
                        "frac": round(alg_bytes(a, b, o) / t / PEAK_HBM, 4),
                        "config": launch_config(ap, f, a, b, o)}
    # the dominant op with IEEE flags tracked (NumPy's default modes): flag
    # variant of the kernel plus a stream synchronisation and a 4-byte read
    with ap.errstate(divide="warn", over="warn", invalid="warn"):
        f64o = outs[0]
        with quiet_gc(torch, lambda: ap.multiply(X64, c64, out=f64o)):
            t0 = time.perf_counter()
            for _ in range(10):
                ap.multiply(X64, c64, out=f64o)
            tt = (time.perf_counter() - t0) / 10
    fpe_tracking = {"op": "mul_f64", "us": round(tt * 1e6, 1), "GB/s": round(alg_bytes(X64, c64, f64o) / tt / 1e9, 1),
                    "note": "per call incl. the flag fetch (host wall clock)"}
    dom = per_op["mul_f64"]
    roofline = {"bound": "hbm", "achieved": dom["GB/s"], "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": dom["frac"], "traffic": None, "kernel": "ew_rows<double,multiply>",
                "algorithmic_bytes_per_launch": alg_bytes(X64, c64, outs[0])}
    traffic_file = os.path.join(HERE, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            roofline["traffic"] = json.load(f).get("mul_f64")

    extra = {"config_policy": ap.get_config_policy(), "policy_settle_steps": settle, "per_op": per_op,
             "fpe_tracking": fpe_tracking}
    del X64, c64, X32, c32, I, j, V32, v32, outs
    torch.cuda.empty_cache()
    if not args.no_extra:
        extra.update(elementwise_extras(ap, torch))
        extra.update(reduction_extras(ap, torch))
        extra["batch"] = batch_extras(ap, torch)
        extra["learned_config"] = None if args.no_sweep else learned_extras(torch)

    cpu = None if args.no_cpu else cpu_baseline()
    line = {"metric": "HBM GB/s (% of peak) ufunc+reduce, bit-exact; batched vs per-call calls/s",
            "value": round(value, 1), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64,f32,i64", "data": "synthetic (seeded NumPy generators)",
            "config": {"workload": "broadcast elementwise: f64/f32 X(8192,8192)*c(8192,1), "
                                   "f32 X.T[::2]*c[::2], i64 I(8192,8192)+j(8192,1)",
                       "bytes_per_step": step_bytes, "frac_of_peak": round(value * 1e9 / PEAK_HBM, 4)},
            "roofline": roofline, "cpu_baseline": cpu, "extra": extra}
    print(json.dumps(line), flush=True)


def settle_policy(torch, step, limit=40):
    """Untimed steps until the advisor has no layout pending or under trial
    (at most `limit`); returns how many ran."""
    from anionpy_amd import costmodel as CM
    for k in range(limit):
        step()
        torch.cuda.synchronize()
        st = CM.stats()
        if st["pending"] == 0 and st["trials_running"] == 0:
            return k + 1
    return limit


def launch_config(ap, f, a, b, o):
    """The launch configuration the call ran with: the advisor's proposal
    (block, vec, blocks_per_cu, split; zeros = the built-in heuristic) or
    "heuristic" when no proposal is cached."""
    from anionpy_amd import costmodel as CM
    from anionpy_amd.batch import binary_call
    got = CM.lookup(binary_call(f.__name__, a, b, out=o))
    return "heuristic" if got is None or not any(got) else list(got)


def timed(torch, fn, reps):
    s = torch.cuda.current_stream()
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with quiet_gc(torch, fn):
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def _rate(nbytes, t):
    return {"us": round(t * 1e6, 1), "GB/s": round(nbytes / t / 1e9, 1), "frac": round(nbytes / t / PEAK_HBM, 4)}


def elementwise_extras(ap, torch):
    """Elementwise and scan workloads beside the step: float64 broadcast add
    (row and column vector), the float64 X.T[::2] view, A + B.T (the
    LDS-transposed operand) in float32 and float64, where(cond, A, column)
    and the float64 -> float32 cast, where(cond, A, B.T), B.T cast into a
    C-ordered float32, float64 A + float32 B.T, and float64 cumsum along
    either axis, all 8192 x 8192."""
    log("elementwise and scan workloads")
    out = {}
    A = ap.asarray(np.random.default_rng(0).uniform(-1, 1, (8192, 8192)))
    row = ap.asarray(np.random.default_rng(9).uniform(-1, 1, 8192))
    col = ap.asarray(np.random.default_rng(10).uniform(-1, 1, (8192, 1)))
    for key, b in (("add_f64_row", row), ("add_f64_col", col)):
        o = ap.add(A, b)
        out[key] = _rate(A.nbytes + b.nbytes + o.nbytes, timed(torch, lambda: ap.add(A, b, out=o), 10))
        del o
    V, v = A.T[::2], col[::2]
    o = ap.multiply(V, v)
    out["view_f64"] = _rate(V.nbytes + v.nbytes + o.nbytes, timed(torch, lambda: ap.multiply(V, v, out=o), 10))
    del o, V, v
    for dt in (np.float32, np.float64):
        Ad = A if dt == np.float64 else ap.asarray(np.random.default_rng(0).uniform(-1, 1, (8192, 8192)).astype(dt))
        Bd = ap.asarray(np.random.default_rng(11).uniform(-1, 1, (8192, 8192)).astype(dt))
        BT = Bd.T
        o = ap.add(Ad, BT)
        out[f"add_T_{np.dtype(dt).name}"] = _rate(Ad.nbytes + BT.nbytes + o.nbytes,
                                                  timed(torch, lambda: ap.add(Ad, BT, out=o), 10))
        # the same call with IEEE flags tracked (NumPy's default modes): the
        # typed LDS tile kernel with flags, plus each call's flag fetch
        with ap.errstate(divide="warn", over="warn", invalid="warn"):
            out[f"add_T_{np.dtype(dt).name}_tracked"] = _rate(Ad.nbytes + BT.nbytes + o.nbytes,
                                                              timed(torch, lambda: ap.add(Ad, BT, out=o), 10))
        del o, Bd, BT, Ad
    cond = ap.asarray(np.random.default_rng(12).random((8192, 8192)) < 0.5)
    o = ap.where(cond, A, col)
    out["where_f64"] = _rate(cond.nbytes + A.nbytes + col.nbytes + o.nbytes,
                             timed(torch, lambda: ap.where(cond, A, col), 10))
    del o
    # a transposed operand through the any-call LDS tile kernel
    B = ap.asarray(np.random.default_rng(13).uniform(-1, 1, (8192, 8192)))
    o = ap.where(cond, A, B.T)
    out["where_f64_T"] = _rate(cond.nbytes + A.nbytes + B.nbytes + o.nbytes,
                               timed(torch, lambda: ap.where(cond, A, B.T), 10))
    del o, cond
    o = A.astype(np.float32)
    out["cast_f64_f32"] = _rate(A.nbytes + o.nbytes, timed(torch, lambda: A.astype(np.float32), 10))
    del o
    o = B.T.astype(np.float32, order="C")
    out["cast_f64_f32_T_to_C"] = _rate(B.nbytes + o.nbytes, timed(torch, lambda: B.T.astype(np.float32, order="C"), 10))
    del o
    Bf = ap.asarray(np.random.default_rng(14).uniform(-1, 1, (8192, 8192)).astype(np.float32))
    o = ap.add(A, Bf.T)
    out["add_f64_f32T"] = _rate(A.nbytes + Bf.nbytes + o.nbytes, timed(torch, lambda: ap.add(A, Bf.T, out=o), 10))
    del o, B, Bf
    for ax in (0, 1):
        o = ap.add.accumulate(A, axis=ax)
        out[f"cumsum_f64_axis{ax}"] = _rate(A.nbytes + o.nbytes,
                                            timed(torch, lambda: ap.add.accumulate(A, axis=ax, out=o), 10))
        del o
    del A, row, col
    torch.cuda.empty_cache()
    return out


def reduction_extras(ap, torch):
    log("reduction workloads")
    out = {}
    x = ap.asarray(np.random.default_rng(4).uniform(-1, 1, 2 ** 28))
    nb = x.nbytes + 8
    t = timed(torch, lambda: ap.add.reduce(x), 10)
    out["sum_f64_2^28"] = {"us": round(t * 1e6, 1), "GB/s": round(nb / t / 1e9, 1), "frac": round(nb / t / PEAK_HBM, 4)}
    t = timed(torch, lambda: ap.maximum.reduce(x), 10)
    out["max_f64_2^28"] = {"us": round(t * 1e6, 1), "GB/s": round(nb / t / 1e9, 1), "frac": round(nb / t / PEAK_HBM, 4)}
    t = timed(torch, lambda: ap.argmax(x), 10)
    out["argmax_f64_2^28"] = {"us": round(t * 1e6, 1), "GB/s": round(nb / t / 1e9, 1),
                              "frac": round(nb / t / PEAK_HBM, 4)}
    del x
    # the same array with NaN at rate 1e-6 (rng(5).random(n) < 1e-6): the
    # maximum / argmax workload as specified, NaN propagation included
    h = np.random.default_rng(4).uniform(-1, 1, 2 ** 28)
    h[np.random.default_rng(5).random(2 ** 28) < 1e-6] = np.nan
    x = ap.asarray(h)
    del h
    for key, fn in (("max_f64_2^28_nan1e-6", lambda: ap.maximum.reduce(x)), ("argmax_f64_2^28_nan1e-6", lambda: ap.argmax(x))):
        t = timed(torch, fn, 10)
        out[key] = {"us": round(t * 1e6, 1), "GB/s": round(nb / t / 1e9, 1), "frac": round(nb / t / PEAK_HBM, 4)}
    del x
    F = ap.asarray(np.random.default_rng(6).uniform(-1, 1, (4096, 65536)).astype(np.float32))
    for ax, nbo in ((0, 65536 * 4), (1, 4096 * 4)):
        nb = F.nbytes + nbo
        t = timed(torch, lambda: ap.add.reduce(F, axis=ax), 5)
        out[f"sum_f32_(4096,65536)_axis{ax}"] = {"us": round(t * 1e6, 1), "GB/s": round(nb / t / 1e9, 1),
                                                 "frac": round(nb / t / PEAK_HBM, 4)}
    del F
    I = ap.asarray(np.random.default_rng(7).integers(-1000, 1000, 2 ** 28, dtype=np.int32))
    nb = I.nbytes + 8
    t = timed(torch, lambda: ap.sum(I), 5)
    out["sum_i32_2^28"] = {"us": round(t * 1e6, 1), "GB/s": round(nb / t / 1e9, 1), "frac": round(nb / t / PEAK_HBM, 4)}
    torch.cuda.empty_cache()
    return out


def batch_extras(ap, torch):
    """The uneven batch: 10,000 calls (seed 5), one persistent launch vs a
    captured hipGraph vs per-call launches; outputs compared bytewise."""
    from anionpy_amd.batch import Batch
    from anionpy_amd.layouts import uneven_batch
    log("uneven batch")
    specs = uneven_batch(seed=5, n=10000)
    t0 = time.perf_counter()
    bt = Batch(specs)
    create_s = time.perf_counter() - t0
    info = bt.info()
    res = {"calls": len(specs), "tiles": info["tiles"], "alg_bytes": info["bytes"],
           "create_ms": round(create_s * 1e3, 2)}
    outs = {}
    for mode in ("per_call", "graph", "persistent"):
        bt.run(mode)  # warm (the graph is captured here)
        torch.cuda.synchronize()
        reps = 3 if mode == "per_call" else 10
        with quiet_gc(torch, lambda: bt.run(mode)):
            t0 = time.perf_counter()
            for _ in range(reps):
                bt.run(mode)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / reps
        res[mode] = {"ms": round(dt * 1e3, 3), "calls_per_s": round(len(specs) / dt, 1),
                     "GB/s": round(info["bytes"] / dt / 1e9, 1)}
        outs[mode] = [s.out._buf.clone() for s in specs]
    same = all(torch.equal(a, b) and torch.equal(a, c)
               for a, b, c in zip(outs["per_call"], outs["graph"], outs["persistent"]))
    res["identical_outputs"] = bool(same)
    res["speedup_persistent_vs_per_call"] = round(res["per_call"]["ms"] / res["persistent"]["ms"], 2)
    res["speedup_graph_vs_per_call"] = round(res["per_call"]["ms"] / res["graph"]["ms"], 2)
    bt.close()
    # the batch autotuner: the persistent run under each tile schedule of
    # BATCH_GRID, then the batch rebuilt with the fastest and timed as above
    log("uneven batch: tile schedule autotuning")
    t0 = time.perf_counter()
    tb = Batch.tuned(specs)
    tune_s = time.perf_counter() - t0
    tb.run("persistent")
    torch.cuda.synchronize()
    with quiet_gc(torch, lambda: tb.run("persistent")):
        t0 = time.perf_counter()
        for _ in range(10):
            tb.run("persistent")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 10
    tuned_same = all(torch.equal(a, s.out._buf) for a, s in zip(outs["persistent"], specs))
    res["persistent_tuned"] = {"ms": round(dt * 1e3, 3), "calls_per_s": round(len(specs) / dt, 1),
                               "GB/s": round(info["bytes"] / dt / 1e9, 1),
                               "schedule": {"ew_tile": tb.schedule[0], "int_seg": tb.schedule[1],
                                            "claim": tb.schedule[2]},
                               "default_schedule_us": round(float(tb.tune_times[0]), 1),
                               "tuned_schedule_us": round(float(np.min(tb.tune_times[tb.tune_times > 0])), 1),
                               "schedules_tried": int(len(tb.tune_times)), "tuning_s": round(tune_s, 2),
                               "identical_outputs": bool(tuned_same)}
    res["speedup_persistent_tuned_vs_per_call"] = round(res["per_call"]["ms"] / res["persistent_tuned"]["ms"], 2)
    tb.close()
    del specs, outs
    torch.cuda.empty_cache()
    return res


def learned_extras(torch):
    """The committed cost model on 200 layouts generated now (seed 909, not
    seen in training; sizes 2^10-2^27) and timed under every grid
    configuration, 5 reps, twice: throughput of the learned choices and of
    the heuristic as a fraction of the exhaustive best, the noise floor (the
    best of one sweep judged by the other), and the stored held-out figures."""
    path = os.path.join(HERE, "anionpy_amd", "costmodel.json")
    if not os.path.exists(path):
        return None
    from anionpy_amd import autotune, costmodel as CM
    from anionpy_amd.layouts import Pool, tuning_layouts
    from anionpy_amd.train_costmodel import MAX_LOG2, MIN_LOG2, POOL_BYTES
    log("learned configuration: live sweep")
    model, meta = CM.load()
    pool = Pool(POOL_BYTES)
    specs = tuning_layouts(909, 200, pool, min_log2=MIN_LOG2, max_log2=MAX_LOG2)
    t = autotune.sweep(specs, reps=5)
    t2 = autotune.sweep(specs, reps=5)
    chosen = model.predict(np.stack([s.features() for s in specs]))
    tt = np.where(t > 0, t, np.inf)
    tb = tt.min(axis=1)

    def total(pick):
        tc = tt[np.arange(len(tt)), pick]
        return float(np.where(np.isfinite(tc), tc, tt[:, 0]).sum())

    tl, th, tn = total(chosen), total(np.zeros(len(specs), dtype=int)), total(CM.best_index(t2))
    ho = meta.get("held_out", {}).get("all", {})
    res = {"live_layouts": len(specs), "exhaustive_us": round(float(tb.sum()), 1), "learned_us": round(tl, 1),
           "heuristic_us": round(th, 1), "noise_floor_us": round(tn, 1),
           "learned_throughput_vs_exhaustive": round(float(tb.sum() / tl), 4),
           "heuristic_throughput_vs_exhaustive": round(float(tb.sum() / th), 4),
           "noise_floor_throughput_vs_exhaustive": round(float(tb.sum() / tn), 4),
           "learned_fallback_share": round(float(np.mean(chosen == 0)), 3),
           "stored_held_out": ho}
    del specs, pool
    torch.cuda.empty_cache()
    return res


def cpu_baseline(min_seconds=10.0):
    """The scalar C oracle (single thread) on the step workload itself, at
    full size, repeated until at least `min_seconds` of CPU work; NumPy on
    the same host and arrays is timed beside it."""
    sys.path.insert(0, HERE)
    import oracle as O
    X = np.random.default_rng(1).standard_normal((8192, 8192))
    c = np.random.default_rng(2).uniform(0.5, 2, (8192, 1))
    X32, c32 = X.astype(np.float32), c.astype(np.float32)
    I = np.random.default_rng(3).integers(-2 ** 62, 2 ** 62, (8192, 8192))
    j = np.random.default_rng(8).integers(-2 ** 62, 2 ** 62, (8192, 1))
    work = [("multiply", X, c), ("multiply", X32, c32), ("multiply", X32.T[::2], c32[::2]), ("add", I, j)]
    nbytes, reps = 0, 0
    t0 = time.perf_counter()
    while reps == 0 or time.perf_counter() - t0 < min_seconds:
        for op, a, b in work:
            o = O.binary(op, a, b)
            nbytes += a.nbytes + b.nbytes + o.nbytes  # X.T[::2]: its own 128 MiB, as on the GPU
            del o
        reps += 1
    dt = time.perf_counter() - t0
    # the same loop on every host core (OpenMP, output rows split per thread)
    mbytes, mreps, threads = 0, 0, 1
    t2 = time.perf_counter()
    while mreps == 0 or time.perf_counter() - t2 < min_seconds / 2:
        for op, a, b in work:
            o, threads = O.binary_omp(op, a, b)
            mbytes += a.nbytes + b.nbytes + o.nbytes
            del o
        mreps += 1
    dmt = time.perf_counter() - t2
    t1 = time.perf_counter()
    for op, a, b in work:
        getattr(np, op)(a, b)
    dnp = time.perf_counter() - t1
    return {"value": round(nbytes / dt / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"C oracle (scalar restatement of the reference loop) over the full step workload, "
                      f"{reps} pass(es), {nbytes / 1e9:.1f} GB algorithmic in {dt:.1f} s; host CPU: {cpu_name()}",
            "openmp": {"value": round(mbytes / dmt / 1e9, 3), "unit": "GB/s", "cores": threads,
                       "sample": f"{mreps} pass(es) of the same workload, output rows split over {threads} threads"},
            "numpy_same_host_GBps": round(nbytes / reps / dnp / 1e9, 3)}


def cpu_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip() + f" x{os.cpu_count()}"
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
