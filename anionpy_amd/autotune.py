"""Exhaustive launch-configuration autotuner.

Every configuration computes the same bytes (the kernels' arithmetic order
never depends on the launch shape), so tuning only trades speed.  The grid
spans the four ionp_config knobs: workgroup size (tile), vector width,
workgroups per CU and the unroll / split factor.  Entry 0 is the built-in
heuristic (all zeros).  Timings come from the native sweep ionp_autotune
(HIP events around every launch, fastest of `reps`).
"""
import ctypes
import itertools

import numpy as np

from . import _call, _lib
from .batch import call_array

BLOCKS = (64, 128, 256)
VECS = (0, 1)          # 0: widest vector the layout allows; 1: scalar accesses
BPCS = (1, 2, 4, 8, 16, 32)
SPLITS = (1, 4)
HEADS = {"block": BLOCKS, "vec": VECS, "blocks_per_cu": BPCS, "split": SPLITS}

GRID = [(0, 0, 0, 0)] + list(itertools.product(BLOCKS, VECS, BPCS, SPLITS))


def config_array(grid=GRID):
    arr = (_lib.Config * len(grid))()
    for i, (b, v, p, s) in enumerate(grid):
        arr[i] = _lib.Config(b, v, p, s)
    return arr


def knob_labels(idx, grid=GRID):
    """Per-head class indices of grid entry idx (entry 0 maps to the defaults
    block 256, vec 0, blocks_per_cu 8, split 4)."""
    b, v, p, s = grid[idx] if idx else (256, 0, 8, 4)
    return [BLOCKS.index(b), VECS.index(v), BPCS.index(p), SPLITS.index(s)]


def config_index(labels, grid=GRID):
    b, v, p, s = (HEADS["block"][labels[0]], HEADS["vec"][labels[1]], HEADS["blocks_per_cu"][labels[2]],
                  HEADS["split"][labels[3]])
    return grid.index((b, v, p, s))


def sweep(specs, reps=3, grid=GRID, chunk=256, refine=0):
    """times[i, j]: microseconds of call i under grid[j] (negative: unusable).

    refine=0: every configuration timed `reps` times (fastest kept).
    refine=k: successive halving; every configuration is timed once, then
    the k fastest of each call and the heuristic (entry 0) are timed `reps`
    more times and keep their fastest sample; the others keep their single
    timing."""
    first_reps = 1 if refine else reps
    cfgs = config_array(grid)
    out = np.empty((len(specs), len(grid)), dtype=np.float32)
    L = _lib.load()
    for s0 in range(0, len(specs), chunk):
        part = specs[s0:s0 + chunk]
        arr = call_array(part)
        t = (ctypes.c_float * (len(part) * len(grid)))()
        _call(L.ionp_autotune, arr, len(part), cfgs, len(grid), first_reps, t, _lib.stream_handle())
        out[s0:s0 + len(part)] = np.frombuffer(t, dtype=np.float32).reshape(len(part), len(grid))
    if refine:
        for i, spec in enumerate(specs):
            row = np.where(out[i] > 0, out[i], np.inf)
            cand = sorted(set(np.argsort(row)[:refine].tolist()) | {0})
            cand = [j for j in cand if np.isfinite(row[j]) or j == 0]
            sub = config_array([grid[j] for j in cand])
            t = (ctypes.c_float * len(cand))()
            _call(L.ionp_autotune, call_array([spec]), 1, sub, len(cand), reps, t, _lib.stream_handle())
            for k, j in enumerate(cand):
                if t[k] > 0 and (out[i, j] <= 0 or t[k] < out[i, j]):
                    out[i, j] = t[k]
    return out


def best(times):
    """Index of the fastest usable configuration per row."""
    t = np.where(times > 0, times, np.inf)
    return np.argmin(t, axis=1)
