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
    c.kind = _lib.CALL_REDUCE
    c.naxes = 1
    c.axes[0] = 0
    assert L.ionp_call_features(ctypes.byref(c), (ctypes.c_float * _lib.NFEATURES)()) == 0


def test_autotune_grid_round_trip():
    from anionpy_amd import autotune
    for i in range(1, len(autotune.GRID)):
        assert autotune.config_index(autotune.knob_labels(i)) == i
    assert autotune.GRID[0] == (0, 0, 0, 0)


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
