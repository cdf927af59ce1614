"""ionp_reduce_noalloc (the reduction entry point that allocates nothing:
its scratch is one fixed array per GPU) against ionp_reduce, called through
the C ABI: two calls in a row per layout, the second smaller, bit for bit;
a reduction needing more scratch than the array holds is rejected with its
output untouched."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _axes(axis):
    if axis is None:
        return None, -1
    axis = (axis,) if isinstance(axis, int) else tuple(axis)
    return (ctypes.c_int32 * len(axis))(*axis), len(axis)


def _run(ap, fn, op, x, axis, dtype=None):
    """fn(op, x, axis) into a fresh output poisoned with 0xAB; returns
    (status, output bytes, message)."""
    L = ap._lib.load()
    axes, naxes = _axes(axis)
    d = x._desc()
    t = ap._lib.ArrayDesc()
    dt = -1 if dtype is None else ap._CODE[np.dtype(dtype)]
    assert L.ionp_plan_reduce(op, ctypes.byref(d), axes, naxes, 0, dt, ctypes.byref(t)) == 0
    out = ap.empty_strided(tuple(t.shape[: t.ndim]), tuple(t.strides[: t.ndim]), ap._NP_DTYPES[t.dtype])
    out._buf.fill_(0xAB)
    do = out._desc()
    rc = fn(op, ctypes.byref(d), axes, naxes, 0, None, ctypes.byref(do), None, ap._lib.stream_handle())
    torch.cuda.synchronize()
    msg = L.ionp_last_error().decode() if rc else ""
    return rc, out._buf.cpu().numpy().tobytes(), msg


CASES = [
    # (name, op, shape, dtype, axis, out dtype, second shape)
    ("f64 one-output sum, in-launch folds", "ADD", (2 ** 22,), np.float64, 0, None, (2 ** 18 + 3,)),
    ("f32 row sums, partials folded after the launch", "ADD", (300, 40000), np.float32, 1, None, (100, 20000)),
    ("f64 narrow chains", "ADD", (5000, 7), np.float64, 0, None, (300, 3)),
    ("f64 maximum, window summaries", "MAXIMUM", (2 ** 21,), np.float64, None, None, (2 ** 16 + 1,)),
    ("int32 sum to int64, segment partials", "ADD", (2 ** 24,), np.int32, None, None, (2 ** 20,)),
    ("f64 product chains", "MULTIPLY", (3, 70001), np.float64, 1, None, (2, 300)),
    ("f16 rows into float32, converted input", "ADD", (1000, 300), np.float16, 1, np.float32, (10, 300)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_reduce(ap, case):
    _, opname, shape, dt, axis, odt, shape2 = case
    L = ap._lib.load()
    op = getattr(ap._lib, opname)
    rng = np.random.default_rng(41)
    for shp in (shape, shape2):  # the second call is smaller
        h = rng.uniform(0.5, 1.5, shp) if opname == "MULTIPLY" else rng.uniform(-1, 1, shp) * 100
        x = ap.asarray(h.astype(dt))
        rc0, want, m0 = _run(ap, L.ionp_reduce, op, x, axis, odt)
        rc1, got, m1 = _run(ap, L.ionp_reduce_noalloc, op, x, axis, odt)
        assert rc0 == 0, m0
        assert rc1 == 0, m1
        assert got == want, (shp, "differs from ionp_reduce")


def test_noalloc_rejects_oversized(ap):
    """2^20 float16 elements into a float32 loop need a 4 MiB converted copy:
    above the array's 2 MiB stage region."""
    L = ap._lib.load()
    x = ap.asarray(np.ones(2 ** 20, np.float16))
    rc, got, msg = _run(ap, L.ionp_reduce_noalloc, ap._lib.ADD, x, None, np.float32)
    assert rc == 5, msg
    assert "ionp_reduce_noalloc" in msg and "maximum 2 MiB" in msg, msg
    assert got == b"\xab" * 4  # output untouched
    # the entry point keeps working after a rejection
    y = ap.asarray(np.arange(1000, dtype=np.float64))
    rc, got, msg = _run(ap, L.ionp_reduce_noalloc, ap._lib.ADD, y, None)
    assert rc == 0 and np.frombuffer(got, np.float64)[0] == 499500.0, msg
