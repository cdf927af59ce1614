"""Learned launch-configuration model: an fp32 MLP (32 layout features ->
3 x 64 hidden ReLU -> one softmax head per ionp_config knob) that proposes
a configuration per call layout.  Its forward, backward and Adam kernels
are the library's own HIP kernels (ionp_mlp_* in include/ionp_hip.h); torch
tensors only hold the parameters in device memory.

Training data: autotune.sweep timings of seeded random layouts.  The target
of a layout is, per knob, a soft distribution over the knob's values: value
v gets weight exp(-(r_v - 1) / TAU), where r_v is the best time among the
grid configurations with that value divided by the layout's best time, so
values within a few percent of the best share the mass; a fallback head
prefers the heuristic unless the knobs' best beats it by FALLBACK.  Each
layout's targets are scaled by layout_weights (longer layouts count more).  `install` loads
the model into the library's advisor (ionp_advisor_set), which then serves
every call made without a configuration; the built-in heuristic (config 0)
remains the fallback, and the model's regret is measured against the
exhaustive sweep.  Whatever configuration it proposes, results are bit-identical:
configurations only move work between threads.
"""
import ctypes
import json
import os

import numpy as np
import torch

from . import _call, _device, _lib
from .autotune import GRID, HEADS, config_index, knob_labels

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "costmodel.json")
SWEEP_DATA = os.path.join(HERE, "costmodel_sweep.npz")
# one head per knob, then the fallback head: value 1 = run the built-in
# heuristic (grid entry 0, whose per-kernel defaults are not one point of
# the knob product), value 0 = run the knobs the other heads pick
KNOB_HEADS = list(HEADS.values())
HEAD_SIZES = [len(v) for v in KNOB_HEADS] + [2]
HEAD_VALUES = [v for knob in KNOB_HEADS for v in knob] + [0, 1]
HEAD_OFF = np.concatenate([[0], np.cumsum(HEAD_SIZES)]).astype(np.int32)
NOUT = int(HEAD_OFF[-1])
# knob labels of grid entries 1.. (the full knob product)
GRID_LABELS = np.array([knob_labels(i) for i in range(1, len(GRID))], dtype=np.int32)
TAU = 0.01       # soft-target temperature, in units of relative slowdown
FALLBACK = 0.03  # the heuristic is preferred unless the knobs' best is this much faster


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class MLP:
    """Parameters and the HIP training / inference loop."""

    def __init__(self, nin=_lib.NFEATURES, hidden=64, depth=3, seed=0):
        g = torch.Generator().manual_seed(seed)
        sizes = [nin] + [hidden] * depth + [NOUT]
        self.sizes = sizes
        dev = _device()
        self.W, self.b = [], []
        for i in range(len(sizes) - 1):
            bound = (6.0 / (sizes[i] + sizes[i + 1])) ** 0.5
            self.W.append(((torch.rand(sizes[i], sizes[i + 1], generator=g) * 2 - 1) * bound).to(dev))
            self.b.append(torch.zeros(sizes[i + 1], device=dev))
        self._adam = None
        self.step = 0

    # ---- parameters -------------------------------------------------------------
    def params(self):
        return [p for pair in zip(self.W, self.b) for p in pair]

    def state(self):
        return {"sizes": self.sizes, "W": [w.cpu().tolist() for w in self.W], "b": [b.cpu().tolist() for b in self.b]}

    @classmethod
    def from_state(cls, st):
        m = cls.__new__(cls)
        m.sizes = st["sizes"]
        dev = _device()
        m.W = [torch.tensor(w, dtype=torch.float32, device=dev) for w in st["W"]]
        m.b = [torch.tensor(b, dtype=torch.float32, device=dev) for b in st["b"]]
        m._adam = None
        m.step = 0
        return m

    # ---- kernels ---------------------------------------------------------------------
    def forward(self, x):
        """Activations of every layer (the last one is the logits)."""
        L = _lib.load()
        s = _lib.stream_handle()
        acts = [x.contiguous()]
        n = x.shape[0]
        for i, (w, b) in enumerate(zip(self.W, self.b)):
            y = torch.empty(n, w.shape[1], device=x.device)
            relu = int(i < len(self.W) - 1)
            _call(L.ionp_mlp_linear_fwd, _ptr(acts[-1]), _ptr(w), _ptr(b), _ptr(y), n, w.shape[0], w.shape[1], relu, s)
            acts.append(y)
        return acts

    def loss_and_grads(self, x, labels):
        """Mean summed-head cross-entropy and the parameter gradients."""
        L = _lib.load()
        s = _lib.stream_handle()
        acts = self.forward(x)
        n = x.shape[0]
        dz = torch.empty_like(acts[-1])
        loss = torch.empty(n + 1, device=x.device)
        off = torch.tensor(HEAD_OFF, dtype=torch.int32, device=x.device)
        xent = L.ionp_mlp_softmax_xent_soft if labels.dtype == torch.float32 else L.ionp_mlp_softmax_xent
        _call(xent, _ptr(acts[-1]), _ptr(labels), _ptr(off), len(HEAD_SIZES), n, NOUT, _ptr(dz), _ptr(loss), s)
        grads_W, grads_b = [None] * len(self.W), [None] * len(self.W)
        dy = dz
        for i in range(len(self.W) - 1, -1, -1):
            w = self.W[i]
            dw = torch.empty_like(w)
            db = torch.empty_like(self.b[i])
            dx = torch.empty_like(acts[i]) if i > 0 else None
            relu = int(i < len(self.W) - 1)
            _call(L.ionp_mlp_linear_bwd, _ptr(acts[i]), _ptr(w), _ptr(acts[i + 1]), _ptr(dy),
                  _ptr(dx) if dx is not None else None, _ptr(dw), _ptr(db), n, w.shape[0], w.shape[1], relu, s)
            grads_W[i], grads_b[i] = dw, db
            dy = dx
        return loss, grads_W, grads_b

    def adam_step(self, grads_W, grads_b, lr=3e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = _lib.load()
        s = _lib.stream_handle()
        if self._adam is None:
            self._adam = [(torch.zeros_like(p), torch.zeros_like(p)) for p in self.params()]
        self.step += 1
        grads = [g for pair in zip(grads_W, grads_b) for g in pair]
        for p, g, (m, v) in zip(self.params(), grads, self._adam):
            _call(L.ionp_mlp_adam, _ptr(p), _ptr(g), _ptr(m), _ptr(v), p.numel(), lr, b1, b2, eps, self.step, s)

    def train(self, feats, labels, epochs=40, batch=512, lr=3e-3, seed=0, log=None):
        """labels: int32 (n, heads) knob indices, or float32 (n, NOUT) soft
        targets (a distribution over each head's values)."""
        dev = _device()
        X = torch.as_tensor(np.asarray(feats, dtype=np.float32), device=dev)
        labels = np.asarray(labels)
        Y = torch.as_tensor(labels.astype(np.float32 if labels.dtype.kind == "f" else np.int32), device=dev)
        rng = np.random.default_rng(seed)
        n = X.shape[0]
        hist = []
        for ep in range(epochs):
            perm = torch.as_tensor(rng.permutation(n), device=dev)
            tot = 0.0
            for s0 in range(0, n, batch):
                idx = perm[s0:s0 + batch]
                loss, gW, gb = self.loss_and_grads(X[idx].contiguous(), Y[idx].contiguous())
                self.adam_step(gW, gb, lr=lr)
                tot += float(loss[0]) * len(idx)
            hist.append(tot / n)
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4f}")
        return hist

    def predict(self, feats):
        """Grid index of the proposed configuration for each feature row."""
        X = torch.as_tensor(np.asarray(feats, dtype=np.float32), device=_device())
        z = self.forward(X)[-1].cpu().numpy()
        return decode(z)


def decode(logits):
    """Grid index proposed by each row of logits: 0 (the heuristic) when
    the fallback head says so, else the knob heads' argmaxes."""
    labels = np.stack([logits[:, HEAD_OFF[h]:HEAD_OFF[h + 1]].argmax(axis=1) for h in range(len(HEAD_SIZES))], 1)
    return np.array([0 if l[-1] == 1 else config_index(list(l[:-1])) for l in labels])


def best_index(times):
    """Grid index of each layout's fastest usable configuration."""
    return np.argmin(np.where(times > 0, times, np.inf), axis=1)


def time_ratios(times):
    """Each layout's times divided by its best (unusable: 100)."""
    t = np.where(times > 0, times, np.inf)
    r = t / t.min(axis=1, keepdims=True)
    return np.where(np.isfinite(r), r, 100.0).astype(np.float32)


def soft_targets(ratios, tau=TAU):
    """(n, NOUT) per-knob target distributions from time ratios (n, grid)."""
    r = np.asarray(ratios, dtype=np.float64)
    rg = r[:, 1:]
    q = np.zeros((len(rg), NOUT), dtype=np.float32)

    def softmin(rv):
        z = -(rv - 1.0) / tau
        z -= z.max(axis=1, keepdims=True)
        e = np.exp(z)
        return e / e.sum(axis=1, keepdims=True)

    for h, size in enumerate(HEAD_SIZES[:-1]):
        rv = np.stack([rg[:, GRID_LABELS[:, h] == v].min(axis=1) for v in range(size)], axis=1)
        q[:, HEAD_OFF[h]:HEAD_OFF[h + 1]] = softmin(rv)
    # fallback head: the knobs' best against the heuristic, which wins ties
    # within FALLBACK
    fb = np.stack([rg.min(axis=1), r[:, 0] - FALLBACK], axis=1)
    q[:, HEAD_OFF[-2]:HEAD_OFF[-1]] = softmin(fb)
    return q


def layout_weights(t_best):
    """Training weight of each layout: 1 + log2(best time / 10 us), at least
    1, normalised to mean 1 -- a layout that takes longer counts more (the
    aggregate throughput is theirs), without drowning the many short ones."""
    w = np.clip(1.0 + np.log2(np.asarray(t_best, dtype=np.float64) / 10.0), 1.0, None)
    return (w / w.mean()).astype(np.float32)


def regret(times, chosen):
    """Aggregate slowdown of `chosen` configs vs the per-layout best:
    sum(t_chosen) / sum(t_best) - 1, and the mean per-layout ratio - 1."""
    t = np.where(times > 0, times, np.inf)
    tb = t.min(axis=1)
    tc = t[np.arange(len(t)), chosen]
    tc = np.where(np.isfinite(tc), tc, t[:, 0])  # an unusable proposal falls back to the heuristic
    return float(tc.sum() / tb.sum() - 1.0), float(np.mean(tc / tb) - 1.0)


def save(model, meta=None, path=WEIGHTS):
    st = model.state()
    st["meta"] = meta or {}
    st["grid"] = GRID
    with open(path, "w") as f:
        json.dump(st, f)


def load(path=WEIGHTS):
    with open(path) as f:
        st = json.load(f)
    return MLP.from_state(st), st.get("meta", {})


def install(model):
    """Load `model` (an MLP) into the library's advisor: from now on every
    call made without a configuration runs the model's proposal for its
    layout (a layout's first call runs the heuristic while the proposal is
    computed on the GPU, without waiting)."""
    _install(model.sizes, [w.detach().cpu().numpy() for w in model.W], [b.detach().cpu().numpy() for b in model.b])


def install_file(path=WEIGHTS):
    """install() from saved weights, without building device tensors."""
    with open(path) as f:
        st = json.load(f)
    _install(st["sizes"], [np.asarray(w, np.float32) for w in st["W"]], [np.asarray(b, np.float32) for b in st["b"]])


def _install(sizes_, Ws, bs):
    sizes = (ctypes.c_int32 * len(sizes_))(*sizes_)
    flat = np.concatenate([np.concatenate([np.ravel(w), np.ravel(b)]) for w, b in zip(Ws, bs)]).astype(np.float32)
    heads = (ctypes.c_int32 * len(HEAD_SIZES))(*HEAD_SIZES)
    values = (ctypes.c_int32 * len(HEAD_VALUES))(*HEAD_VALUES)
    _call(_lib.load().ionp_advisor_set, sizes, len(sizes_) - 1, flat.ctypes.data_as(ctypes.c_void_p), heads,
          len(HEAD_SIZES), values)


def uninstall():
    """Every call made without a configuration runs the heuristic."""
    _call(_lib.load().ionp_advisor_set, None, 0, None, None, 0, None)


def prime(specs):
    """Evaluate the installed model for these calls' layouts now (batched
    passes, one wait), so their first calls already run the learned
    configuration."""
    from .batch import call_array
    _call(_lib.load().ionp_advisor_prime, call_array(specs), len(specs), _lib.stream_handle())


def stats():
    """Counters of the advisor: calls that ran its configuration, calls that
    ran the heuristic while their proposal was pending or under trial,
    layouts evaluated, layouts pending, and the trials' outcomes (each
    proposal is timed against the heuristic on live calls and the faster
    one kept)."""
    v = (ctypes.c_int64 * 7)()
    _call(_lib.load().ionp_advisor_stats, v)
    return {"learned_calls": v[0], "heuristic_calls": v[1], "layouts_evaluated": v[2], "pending": v[3],
            "trials_kept_proposal": v[4], "trials_kept_heuristic": v[5], "trials_running": v[6]}


def lookup(spec):
    """The advisor's cached configuration for a call's layout, or None."""
    cfg, found = _lib.Config(), ctypes.c_int32()
    _call(_lib.load().ionp_advisor_lookup, ctypes.byref(spec.rec), ctypes.byref(cfg), ctypes.byref(found))
    return (cfg.block, cfg.vec, cfg.blocks_per_cu, cfg.split) if found.value else None
