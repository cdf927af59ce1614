"""Learned launch-configuration model: an fp32 MLP (24 layout features ->
3 x 64 hidden ReLU -> one softmax head per ionp_config knob) that proposes
a configuration per call.  Its forward, backward and Adam kernels are the
library's own HIP kernels (ionp_mlp_* in include/ionp_hip.h); torch tensors
only hold the parameters in device memory.

Training data: autotune.sweep timings of seeded random layouts; the label
of a layout is the knob setting of its fastest configuration.  The built-in
heuristic (config 0) remains the fallback, and the exhaustive sweep is the
judge of the model's regret.  Whatever configuration it proposes, results
are bit-identical: configurations only move work between threads.
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
HEAD_SIZES = [len(v) for v in HEADS.values()]
HEAD_OFF = np.concatenate([[0], np.cumsum(HEAD_SIZES)]).astype(np.int32)
NOUT = int(HEAD_OFF[-1])


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
        _call(L.ionp_mlp_softmax_xent, _ptr(acts[-1]), _ptr(labels), _ptr(off), len(HEAD_SIZES), n, NOUT, _ptr(dz),
              _ptr(loss), s)
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
        dev = _device()
        X = torch.as_tensor(np.asarray(feats, dtype=np.float32), device=dev)
        Y = torch.as_tensor(np.asarray(labels, dtype=np.int32), device=dev)
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
        labels = np.stack([z[:, HEAD_OFF[h]:HEAD_OFF[h + 1]].argmax(axis=1) for h in range(len(HEAD_SIZES))], 1)
        return np.array([config_index(list(l)) for l in labels])


def labels_of(times):
    """Knob labels of each layout's fastest configuration."""
    t = np.where(times > 0, times, np.inf)
    return np.array([knob_labels(int(i)) for i in np.argmin(t, axis=1)], dtype=np.int32)


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


class Advisor:
    """Per-call launch configuration proposed by the model, cached by the
    call's layout features (ionp_call_features): the MLP's HIP forward runs
    once per distinct layout.  Grid entry 0 is the built-in heuristic."""

    def __init__(self, model):
        self.model = model
        self.cache = {}

    def config(self, rec):
        f = (ctypes.c_float * _lib.NFEATURES)()
        _call(_lib.load().ionp_call_features, ctypes.byref(rec), f)
        key = bytes(f)
        cfg = self.cache.get(key)
        if cfg is None:
            idx = int(self.model.predict(np.frombuffer(f, dtype=np.float32)[None])[0])
            cfg = _lib.Config(*GRID[idx])
            self.cache[key] = cfg
        return cfg

