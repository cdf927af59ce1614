"""Train the launch-configuration model on exhaustive autotuner sweeps.

    python -m anionpy_amd.train_costmodel [--train 50000] [--test 5000]

1. seeded random layouts (anionpy_amd.layouts.tuning_layouts): seed 6 for
   training, seed 7 for the held-out set;
2. every layout timed under every grid configuration (autotune.sweep);
3. the MLP trained on the GPU by the library's HIP kernels (costmodel.MLP);
4. held-out regret of the learned choice vs the exhaustive best, beside the
   built-in heuristic's; weights saved to anionpy_amd/costmodel.json and a
   summary printed as one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np


def main():
    ap_ = argparse.ArgumentParser()
    ap_.add_argument("--train", type=int, default=50000)
    ap_.add_argument("--test", type=int, default=5000)
    ap_.add_argument("--reps", type=int, default=2)
    ap_.add_argument("--epochs", type=int, default=40)
    ap_.add_argument("--out", default=None)
    args = ap_.parse_args()
    import torch
    from . import autotune, costmodel as CM
    from .layouts import Pool, tuning_layouts

    torch.cuda.set_device(0)
    pool = Pool(1 << 29)
    t0 = time.time()
    feats, times = {}, {}
    for name, seed, n in (("train", 6, args.train), ("test", 7, args.test)):
        f_all, t_all = [], []
        for s0 in range(0, n, 2000):  # chunks keep the live call list small
            specs = tuning_layouts(seed * 1000003 + s0, min(2000, n - s0), pool)
            f_all.append(np.stack([s.features() for s in specs]))
            t_all.append(autotune.sweep(specs, reps=args.reps))
            print(f"{name}: {s0 + len(specs)}/{n} layouts swept ({time.time() - t0:.0f} s)", file=sys.stderr, flush=True)
            del specs
        feats[name], times[name] = np.concatenate(f_all), np.concatenate(t_all)
    sweep_s = time.time() - t0
    m = CM.MLP(seed=0)
    t1 = time.time()
    hist = m.train(feats["train"], CM.labels_of(times["train"]), epochs=args.epochs, batch=512,
                   log=lambda s: print(s, file=sys.stderr, flush=True))
    train_s = time.time() - t1
    chosen = m.predict(feats["test"])
    learned = CM.regret(times["test"], chosen)
    heuristic = CM.regret(times["test"], np.zeros(len(chosen), dtype=int))
    tb = np.where(times["test"] > 0, times["test"], np.inf).min(axis=1)
    meta = {"train_layouts": int(len(feats["train"])), "test_layouts": int(len(feats["test"])),
            "grid": len(autotune.GRID), "reps": args.reps, "epochs": args.epochs,
            "loss_first": hist[0], "loss_last": hist[-1],
            "regret_learned_aggregate": learned[0], "regret_learned_mean": learned[1],
            "regret_heuristic_aggregate": heuristic[0], "regret_heuristic_mean": heuristic[1],
            "best_total_us": float(tb.sum()), "sweep_seconds": round(sweep_s, 1), "train_seconds": round(train_s, 1)}
    CM.save(m, meta)
    if args.out:
        np.savez_compressed(args.out, f_train=feats["train"], t_train=times["train"], f_test=feats["test"],
                            t_test=times["test"])
    print(json.dumps(meta), flush=True)


if __name__ == "__main__":
    main()
