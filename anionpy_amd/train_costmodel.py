"""Train the launch-configuration model on exhaustive autotuner sweeps.

    python -m anionpy_amd.train_costmodel [--train 50000] [--test 5000] [--save-data PATH]
    python -m anionpy_amd.train_costmodel --data anionpy_amd/costmodel_sweep.npz

1. seeded random layouts (anionpy_amd.layouts.tuning_layouts): seed 6 for
   training, seed 7 for the held-out set;
2. every layout timed under every grid configuration (autotune.sweep), or
   a stored sweep loaded with --data (anionpy_amd/costmodel_sweep.npz holds
   the 50k / 5k sweep measured on an MI355X: training features and the
   index of each layout's fastest configuration, held-out features and
   full timings);
3. the MLP trained on the GPU by the library's HIP kernels (costmodel.MLP);
4. held-out regret of the learned choice vs the exhaustive best, beside the
   built-in heuristic's; weights saved (--weights, default
   anionpy_amd/costmodel.json) and a summary printed as one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np


def sweep_layouts(n_train, n_test, reps):
    """Features and grid timings of freshly generated layouts."""
    from . import autotune
    from .layouts import Pool, tuning_layouts
    pool = Pool(1 << 29)
    t0 = time.time()
    feats, times = {}, {}
    for name, seed, n in (("train", 6, n_train), ("test", 7, n_test)):
        f_all, t_all = [], []
        for s0 in range(0, n, 2000):  # chunks keep the live call list small
            specs = tuning_layouts(seed * 1000003 + s0, min(2000, n - s0), pool)
            f_all.append(np.stack([s.features() for s in specs]))
            t_all.append(autotune.sweep(specs, reps=reps))
            print(f"{name}: {s0 + len(specs)}/{n} layouts swept ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
            del specs
        feats[name], times[name] = np.concatenate(f_all), np.concatenate(t_all)
    return feats, times


def load_data(path):
    """(f_train, best_train, f_test, t_test) from a stored sweep."""
    d = np.load(path)
    return d["f_train"], d["best_train"].astype(np.int64), d["f_test"], d["t_test"]


def save_data(path, f_train, t_train, f_test, t_test):
    from .autotune import best
    np.savez_compressed(path, f_train=f_train, best_train=best(t_train).astype(np.int8), f_test=f_test, t_test=t_test)


def fit(f_train, best_train, f_test, t_test, epochs=40, batch=512, seed=0, log=None):
    """Train on the GPU and score on the held-out timings.  Returns the
    model and a summary dict."""
    from . import costmodel as CM
    from .autotune import knob_labels
    labels = np.array([knob_labels(int(i)) for i in best_train], dtype=np.int32)
    m = CM.MLP(seed=seed)
    t1 = time.time()
    hist = m.train(f_train, labels, epochs=epochs, batch=batch, seed=seed, log=log)
    train_s = time.time() - t1
    chosen = m.predict(f_test)
    learned = CM.regret(t_test, chosen)
    heuristic = CM.regret(t_test, np.zeros(len(chosen), dtype=int))
    tb = np.where(t_test > 0, t_test, np.inf).min(axis=1)
    meta = {"train_layouts": int(len(f_train)), "test_layouts": int(len(f_test)), "epochs": epochs,
            "loss_first": hist[0], "loss_last": hist[-1],
            "regret_learned_aggregate": learned[0], "regret_learned_mean": learned[1],
            "regret_heuristic_aggregate": heuristic[0], "regret_heuristic_mean": heuristic[1],
            "best_total_us": float(tb.sum()), "train_seconds": round(train_s, 1)}
    return m, meta


def main():
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--train", type=int, default=50000)
    ap_.add_argument("--test", type=int, default=5000)
    ap_.add_argument("--reps", type=int, default=2)
    ap_.add_argument("--epochs", type=int, default=40)
    ap_.add_argument("--data", default=None, help="stored sweep to train from instead of sweeping")
    ap_.add_argument("--save-data", default=None, help="write the sweep in the stored-sweep format")
    ap_.add_argument("--weights", default=None, help="where to write the trained weights")
    args = ap_.parse_args()
    import torch
    from . import autotune, costmodel as CM

    torch.cuda.set_device(0)
    t0 = time.time()
    if args.data:
        f_train, best_train, f_test, t_test = load_data(args.data)
        sweep_s = 0.0
    else:
        feats, times = sweep_layouts(args.train, args.test, args.reps)
        sweep_s = time.time() - t0
        f_train, f_test, t_test = feats["train"], feats["test"], times["test"]
        best_train = autotune.best(times["train"])
        if args.save_data:
            save_data(args.save_data, f_train, times["train"], f_test, t_test)
    m, meta = fit(f_train, best_train, f_test, t_test, epochs=args.epochs,
                  log=lambda s: print(s, file=sys.stderr, flush=True))
    meta.update({"grid": len(autotune.GRID), "reps": args.reps, "sweep_seconds": round(sweep_s, 1),
                 "data": args.data or "live sweep"})
    CM.save(m, meta, path=args.weights or CM.WEIGHTS)
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
