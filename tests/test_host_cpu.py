"""Host-side pieces that need no GPU: layout features, the autotune grid and
the float64 MLP reference (checked by finite differences)."""
import ctypes

import numpy as np

import oracle.mlp_ref as R


def _desc(shape, strides, dtype_code):
    from anionpy_amd import _lib
    d = _lib.ArrayDesc()
    d.dtype = dtype_code
    d.ndim = len(shape)
    for i, (n, s) in enumerate(zip(shape, strides)):
        d.shape[i] = n
        d.strides[i] = s
    return d


def test_call_features_shape_and_classes():
    from anionpy_amd import _lib
    L = _lib.load()
    c = _lib.Call()
    c.kind = _lib.CALL_BINARY
    c.op = _lib.ADD
    c.inp[0] = _desc((64, 128), (128, 1), _lib.FLOAT64)
    c.inp[1] = _desc((64, 1), (1, 0), _lib.FLOAT64)
    c.out = _desc((64, 128), (128, 1), _lib.FLOAT64)
    f = (ctypes.c_float * _lib.NFEATURES)()
    assert L.ionp_call_features(ctypes.byref(c), f) == 0
    f = np.frombuffer(f, dtype=np.float32)
    assert f[0] == 0.5  # two coalesced dims
    assert f[6] == 1.0 and f[10] == 1.0  # input 0 contiguous inner, input 1 broadcast inner
    assert f[14] == 1.0 and f[17] == 1.0  # float64, arithmetic
    assert f[24] == 1.0  # float64 output
    assert f[25] == 0.0 and f[26] == 0.0  # no negative inner stride
    assert abs(f[28] - np.log2(2.0) / 24) < 1e-7  # largest inner stride 1
    # bytes: a 64 KiB + c 512 B + out 64 KiB
    assert abs(f[29] - np.log2(64 * 128 * 8 * 2 + 64 * 8 + 1) / 36) < 1e-6
    assert f[30] == 0.0 and f[31] == 0.0
    c.inp[0] = _desc((64, 128), (128, -1), _lib.FLOAT64)  # reversed inner axis
    g = (ctypes.c_float * _lib.NFEATURES)()
    assert L.ionp_call_features(ctypes.byref(c), g) == 0
    assert np.frombuffer(g, dtype=np.float32)[25] == 1.0
    c.kind = _lib.CALL_REDUCE
    c.naxes = 1
    c.axes[0] = 0
    assert L.ionp_call_features(ctypes.byref(c), (ctypes.c_float * _lib.NFEATURES)()) == 0


def test_autotune_grid_round_trip():
    from anionpy_amd import autotune
    for i in range(1, len(autotune.GRID)):
        assert autotune.config_index(autotune.knob_labels(i)) == i
    assert autotune.GRID[0] == (0, 0, 0, 0)


def _finite_differences(lab, Ws, bs, off, x):
    _, gW, gb = R.loss_and_grads(x, lab, Ws, bs, off)
    eps = 1e-6
    for which, params, grads in (("W", Ws, gW), ("b", bs, gb)):
        for k in range(2):
            p = params[k]
            it = np.nditer(p, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = p[idx]
                p[idx] = old + eps
                lp, _, _ = R.loss_and_grads(x, lab, Ws, bs, off)
                p[idx] = old - eps
                lm, _, _ = R.loss_and_grads(x, lab, Ws, bs, off)
                p[idx] = old
                assert abs((lp - lm) / (2 * eps) - grads[k][idx]) < 1e-6


def test_mlp_reference_soft_targets():
    """Soft-target loss: finite differences, and one-hot targets give the
    hard-label loss and gradients exactly."""
    rng = np.random.default_rng(2)
    sizes = [5, 7, 6]
    Ws = [rng.normal(size=(sizes[i], sizes[i + 1])) for i in range(2)]
    bs = [rng.normal(size=sizes[i + 1]) for i in range(2)]
    off = np.array([0, 2, 6])
    x = rng.normal(size=(4, 5))
    q = np.zeros((4, 6))
    for h in range(2):
        w = rng.uniform(0.1, 1, (4, off[h + 1] - off[h]))
        q[:, off[h]:off[h + 1]] = w / w.sum(1, keepdims=True)
    _finite_differences(q, Ws, bs, off, x)
    # weighted samples: a row scaled by m counts m times
    m = rng.uniform(0.2, 3.0, (4, 1))
    _finite_differences(q * m, Ws, bs, off, x)
    l1, _, _ = R.loss_and_grads(x, q * m, Ws, bs, off)
    rows = [R.loss_and_grads(x[i:i + 1], q[i:i + 1], Ws, bs, off)[0] for i in range(4)]
    assert abs(l1 - sum(m[i, 0] * rows[i] for i in range(4)) / 4) < 1e-12
    lab = np.stack([rng.integers(0, 2, 4), rng.integers(0, 4, 4)], 1)
    onehot = np.zeros((4, 6))
    for h in range(2):
        onehot[np.arange(4), off[h] + lab[:, h]] = 1.0
    lh, gWh, gbh = R.loss_and_grads(x, lab, Ws, bs, off)
    ls, gWs, gbs = R.loss_and_grads(x, onehot, Ws, bs, off)
    assert abs(lh - ls) < 1e-12
    for a, b in zip(gWh + gbh, gWs + gbs):
        assert np.abs(a - b).max() < 1e-12


def test_soft_targets_follow_the_timings():
    """Per knob, the value whose best configuration is fastest gets the most
    mass; values far slower get ~none; each head sums to 1."""
    from anionpy_amd import costmodel as CM
    from anionpy_amd.autotune import GRID, knob_labels
    rng = np.random.default_rng(4)
    t = rng.uniform(2.0, 3.0, (3, len(GRID))).astype(np.float32)
    t[:, 0] = 2.5
    best = [5, 40, 71]
    for i, j in enumerate(best):
        t[i, j] = 1.0
    q = CM.soft_targets(CM.time_ratios(t))
    for i, j in enumerate(best):
        lab = knob_labels(j) + [0]  # the fallback head: the knobs (the heuristic is 2.5x slower)
        for h in range(len(CM.HEAD_SIZES)):
            seg = q[i, CM.HEAD_OFF[h]:CM.HEAD_OFF[h + 1]]
            assert abs(seg.sum() - 1.0) < 1e-5
            assert seg.argmax() == lab[h]
            assert seg[lab[h]] > 0.999  # every other value is >= 2x slower
    # the heuristic within FALLBACK of the best: the fallback head picks it
    t3 = t.copy()
    t3[1, 0] = 1.0 + CM.FALLBACK / 2
    fb = CM.soft_targets(CM.time_ratios(t3))[1, CM.HEAD_OFF[-2]:CM.HEAD_OFF[-1]]
    assert fb.argmax() == 1
    z = np.zeros((1, CM.NOUT), dtype=np.float32)
    z[0, CM.HEAD_OFF[-2] + 1] = 1.0
    assert CM.decode(z)[0] == 0
    # an unusable configuration (negative time) never attracts mass
    t2 = t.copy()
    t2[0, 5] = -1.0
    q2 = CM.soft_targets(CM.time_ratios(t2))
    assert CM.best_index(t2)[0] != 5 and np.isfinite(q2).all()


def test_mlp_reference_finite_differences():
    rng = np.random.default_rng(1)
    sizes = [5, 7, 6]
    Ws = [rng.normal(size=(sizes[i], sizes[i + 1])) for i in range(2)]
    bs = [rng.normal(size=sizes[i + 1]) for i in range(2)]
    off = np.array([0, 2, 6])
    x = rng.normal(size=(4, 5))
    lab = np.stack([rng.integers(0, 2, 4), rng.integers(0, 4, 4)], 1)
    _, gW, gb = R.loss_and_grads(x, lab, Ws, bs, off)
    eps = 1e-6
    for which, params, grads in (("W", Ws, gW), ("b", bs, gb)):
        for k in range(2):
            p = params[k]
            it = np.nditer(p, flags=["multi_index"])
            for _ in it:
                idx = it.multi_index
                old = p[idx]
                p[idx] = old + eps
                lp, _, _ = R.loss_and_grads(x, lab, Ws, bs, off)
                p[idx] = old - eps
                lm, _, _ = R.loss_and_grads(x, lab, Ws, bs, off)
                p[idx] = old
                assert abs((lp - lm) / (2 * eps) - grads[k][idx]) < 1e-6


def test_out_overlap_check():
    """The out= overlap check (NumPy computes into a temporary when out
    shares memory with an input other than as the very same view), on
    host byte buffers: separate buffers, the same view, a shifted view, a
    disjoint part of one buffer and a transposed view of the same bytes."""
    import torch
    import anionpy_amd as ap
    f8 = np.dtype(np.float64)
    a_buf = torch.zeros(8 * 100, dtype=torch.uint8)
    o_buf = torch.zeros(8 * 100, dtype=torch.uint8)
    a = ap.ndarray(a_buf, f8, (10, 10), (10, 1))
    assert not ap._needs_temp(ap.ndarray(o_buf, f8, (10, 10), (10, 1)), [a])
    assert not ap._needs_temp(ap.ndarray(a_buf, f8, (10, 10), (10, 1)), [a])
    assert ap._needs_temp(ap.ndarray(a_buf, f8, (9, 10), (10, 1), 1), [ap.ndarray(a_buf, f8, (9, 10), (10, 1))])
    half = ap.ndarray(a_buf, f8, (5, 10), (10, 1))
    assert not ap._needs_temp(ap.ndarray(a_buf, f8, (5, 10), (10, 1), 50), [half])
    assert ap._needs_temp(ap.ndarray(a_buf, f8, (10, 10), (1, 10)), [a])
    # a view of the same bytes held by another tensor object
    assert ap._needs_temp(ap.ndarray(a_buf[8:], f8, (9, 10), (10, 1)), [ap.ndarray(a_buf, f8, (9, 10), (10, 1))])
