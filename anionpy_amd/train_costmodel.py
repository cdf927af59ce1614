"""Train the launch-configuration model on exhaustive autotuner sweeps.

    # sweep (on the GPU), in parts that each fit one session
    python -m anionpy_amd.train_costmodel sweep --set train --blocks 0:13 --out sweep/
    python -m anionpy_amd.train_costmodel sweep --set test --block 500 --blocks 0:10 --out sweep/
    # pack the blocks into the stored sweep (CPU), then train on the GPU from
    # it and score on the held-out set
    python -m anionpy_amd.train_costmodel pack --train sweep/train_*.npz --test sweep/test_*.npz \
        --out anionpy_amd/costmodel_sweep.npz
    python -m anionpy_amd.train_costmodel fit --data anionpy_amd/costmodel_sweep.npz

1. seeded random layouts (anionpy_amd.layouts.tuning_layouts), in blocks of
   B layouts (2,000 for training, 500 for the held-out set): block k has
   seed 6 * 1000003 + B k (training) or 7 * 1000003 + B k; sizes log-uniform in
   [2^10, 2^27] elements (the heuristic's losses concentrate above 2^21);
2. training layouts: every grid configuration timed once, then the 8
   fastest and the heuristic 5 more times each (autotune.sweep, refine=8);
   held-out layouts: every configuration timed 5 times, in two independent
   sweeps, so that run-to-run noise can be reported beside each regret;
3. the MLP trained on the GPU by the library's HIP kernels (costmodel.MLP)
   on soft per-knob targets (costmodel.soft_targets) weighted by each
   layout's time (costmodel.layout_weights);
4. held-out regret of the learned choice vs the exhaustive best, beside the
   built-in heuristic's and the noise floor (the best of one held-out sweep
   judged by the other), overall and per op class; weights saved (--weights,
   default anionpy_amd/costmodel.json) and a summary printed as one JSON line.

anionpy_amd/costmodel_sweep.npz holds the stored sweep measured on an MI355X
(training features and timing ratios, held-out features, tags and both
timing sweeps); `fit --data` trains from it.
"""
import argparse
import json
import sys
import time

import numpy as np

BLOCK = 2000
SEEDS = {"train": 6, "test": 7}
MIN_LOG2, MAX_LOG2 = 10, 27
POOL_BYTES = 8 << 30


def sweep_blocks(which, b0, b1, outdir, block=BLOCK, reps=5, refine=8, log=None):
    """Sweep layout blocks [b0, b1) of a set, writing each block to
    outdir/<set>_<block>.npz (features, tags, timings) as soon as it is done.
    The held-out set is swept twice with every configuration at `reps`."""
    import os
    from . import autotune
    from .layouts import Pool, tuning_layouts
    pool = Pool(POOL_BYTES)
    t0 = time.time()
    os.makedirs(outdir, exist_ok=True)
    for b in range(b0, b1):
        specs = tuning_layouts(SEEDS[which] * 1000003 + block * b, block, pool, min_log2=MIN_LOG2, max_log2=MAX_LOG2)
        d = {"f": np.stack([s.features() for s in specs]), "tags": np.array(["|".join(map(str, s.tag)) for s in specs])}
        if which == "train":
            d["t"] = autotune.sweep(specs, reps=reps, refine=refine)
        else:
            d["t"] = autotune.sweep(specs, reps=reps)
            d["t2"] = autotune.sweep(specs, reps=reps)
        np.savez_compressed(os.path.join(outdir, f"{which}_{b:03d}.npz"), **d)
        if log:
            log(f"{which} block {b}: {len(specs)} layouts swept ({time.time() - t0:.0f} s)")
        del specs


def load_parts(paths):
    parts = [np.load(p) for p in paths]
    out = {k: np.concatenate([d[k] for d in parts]) for k in parts[0].files}
    return out


def load_data(path):
    """(f_train, r_train, tb_train, f_test, t_test, t_test2, tags_test) from
    the stored sweep; r_train holds each training layout's times divided by
    its best, tb_train that best (us)."""
    d = np.load(path)
    return (d["f_train"], d["r_train"].astype(np.float32), d["tb_train"], d["f_test"], d["t_test"], d["t_test2"],
            d["tags_test"])


def save_data(path, f_train, t_train, f_test, t_test, t_test2, tags_test):
    from .costmodel import time_ratios
    tb = np.where(t_train > 0, t_train, np.inf).min(axis=1).astype(np.float32)
    np.savez_compressed(path, f_train=f_train, r_train=time_ratios(t_train).astype(np.float16), tb_train=tb,
                        f_test=f_test, t_test=t_test, t_test2=t_test2, tags_test=tags_test)


def score(model, f_test, t_test, t_test2, tags_test):
    """Regret summary: learned, heuristic and the noise floor, per op class."""
    from . import costmodel as CM
    chosen = model.predict(f_test)
    noise_pick = CM.best_index(t_test2)
    zero = np.zeros(len(chosen), dtype=int)
    classes = np.array([t.split("|")[0] for t in tags_test])
    out = {"all": {"layouts": int(len(f_test))}}
    for name, mask in [("all", np.ones(len(classes), bool))] + [(c, classes == c) for c in sorted(set(classes))]:
        r = out.setdefault(name, {"layouts": int(mask.sum())})
        for key, pick in (("learned", chosen), ("heuristic", zero), ("noise", noise_pick)):
            agg, mean = CM.regret(t_test[mask], pick[mask])
            r[f"regret_{key}_aggregate"] = round(agg, 5)
            r[f"regret_{key}_mean"] = round(mean, 5)
        r["best_total_us"] = round(float(np.where(t_test[mask] > 0, t_test[mask], np.inf).min(axis=1).sum()), 1)
    return out


def fit(f_train, r_train, tb_train, epochs=60, batch=512, seed=0, log=None):
    """Train on the GPU; returns the model and its loss history."""
    from . import costmodel as CM
    targets = CM.soft_targets(r_train) * CM.layout_weights(tb_train)[:, None]
    m = CM.MLP(seed=seed)
    hist = m.train(f_train, targets, epochs=epochs, batch=batch, seed=seed, log=log)
    return m, hist


def main():
    ap_ = argparse.ArgumentParser()
    sub = ap_.add_subparsers(dest="cmd", required=True)
    sw = sub.add_parser("sweep")
    sw.add_argument("--set", choices=("train", "test"), required=True)
    sw.add_argument("--blocks", required=True, help="b0:b1, blocks of 2,000 layouts")
    sw.add_argument("--reps", type=int, default=5)
    sw.add_argument("--block", type=int, default=BLOCK, help="layouts per block")
    sw.add_argument("--out", required=True, help="directory for the per-block files")
    pk = sub.add_parser("pack", help="write swept blocks in the stored-sweep format (no GPU needed)")
    pk.add_argument("--train", nargs="+", required=True)
    pk.add_argument("--test", nargs="+", required=True)
    pk.add_argument("--out", required=True)
    ft = sub.add_parser("fit")
    ft.add_argument("--train", nargs="*", default=[])
    ft.add_argument("--test", nargs="*", default=[])
    ft.add_argument("--data", default=None, help="stored sweep (costmodel_sweep.npz format)")
    ft.add_argument("--save-data", default=None, help="write the parts in the stored-sweep format")
    ft.add_argument("--epochs", type=int, default=60)
    ft.add_argument("--weights", default=None, help="where to write the trained weights")
    args = ap_.parse_args()
    if args.cmd == "pack":
        tr, te = load_parts(args.train), load_parts(args.test)
        save_data(args.out, tr["f"], tr["t"], te["f"], te["t"], te["t2"], te["tags"])
        print(json.dumps({"train_layouts": int(len(tr["f"])), "test_layouts": int(len(te["f"])), "out": args.out}))
        return
    import torch
    torch.cuda.set_device(0)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    from . import costmodel as CM
    from .autotune import GRID
    if args.cmd == "sweep":
        b0, b1 = (int(v) for v in args.blocks.split(":"))
        sweep_blocks(args.set, b0, b1, args.out, block=args.block, reps=args.reps, log=log)
        print(json.dumps({"set": args.set, "blocks": args.blocks}), flush=True)
        return
    if args.data:
        f_train, r_train, tb_train, f_test, t_test, t_test2, tags_test = load_data(args.data)
    else:
        tr, te = load_parts(args.train), load_parts(args.test)
        f_train, r_train = tr["f"], CM.time_ratios(tr["t"])
        tb_train = np.where(tr["t"] > 0, tr["t"], np.inf).min(axis=1)
        f_test, t_test, t_test2, tags_test = te["f"], te["t"], te["t2"], te["tags"]
        if args.save_data:
            save_data(args.save_data, f_train, tr["t"], f_test, t_test, t_test2, tags_test)
    t0 = time.time()
    m, hist = fit(f_train, r_train, tb_train, epochs=args.epochs, log=log)
    meta = {"train_layouts": int(len(f_train)), "test_layouts": int(len(f_test)), "epochs": args.epochs,
            "loss_first": hist[0], "loss_last": hist[-1], "train_seconds": round(time.time() - t0, 1),
            "grid": len(GRID), "sizes_log2": [MIN_LOG2, MAX_LOG2], "reps": 5,
            "held_out": score(m, f_test, t_test, t_test2, tags_test)}
    CM.save(m, meta, path=args.weights or CM.WEIGHTS)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
