"""ufunc keyword semantics on the GPU path: out= aliasing, where=, order=,
reduce out= overlapping its input.  Arithmetic is checked against the
oracle; NumPy indexing only moves the oracle's values into place."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _same_bytes(got, want):
    assert got.dtype == want.dtype and got.shape == want.shape
    assert np.ascontiguousarray(got).tobytes() == np.ascontiguousarray(want).tobytes()


def test_out_overlapping_shifted_view(ap):
    """add(x[1:], x[:-1], out=x[1:]): every input is read before any output
    element is written (NumPy buffers overlapping operands)."""
    x = np.random.default_rng(1).standard_normal(10000)
    want = x.copy()
    want[1:] = O.binary("add", x[1:], x[:-1])
    X = ap.asarray(x)
    ap.add(X[1:], X[:-1], out=X[1:])
    _same_bytes(X.get(), want)


def test_out_same_view_in_place(ap):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((300, 400))
    b = rng.standard_normal(400)
    want = O.binary("multiply", a, b)
    A = ap.asarray(a)
    ap.multiply(A, ap.asarray(b), out=A)
    _same_bytes(A.get(), want)


def test_out_overlapping_transpose(ap):
    """out is the transpose of an input: same memory, different view."""
    a = np.random.default_rng(3).standard_normal((256, 256))
    want = O.binary("subtract", a, a.T)
    A = ap.asarray(a)
    ap.subtract(A, A.T, out=A)
    _same_bytes(A.get(), want)


def test_where_mask_keeps_out(ap):
    rng = np.random.default_rng(4)
    a = rng.standard_normal((70, 90))
    b = rng.standard_normal((70, 1))
    mask = rng.random(90) < 0.5
    init = rng.standard_normal((70, 90))
    want = init.copy()
    full = O.binary("multiply", a, b)
    m2 = np.broadcast_to(mask, (70, 90))
    want[m2] = full[m2]
    out = ap.asarray(init)
    r = ap.multiply(ap.asarray(a), ap.asarray(b), out=out, where=ap.asarray(mask))
    assert r is out
    _same_bytes(out.get(), want)


@pytest.mark.parametrize("order", ["C", "F", "A", "K"])
def test_order_layouts(ap, order):
    rng = np.random.default_rng(5)
    a = np.asfortranarray(rng.standard_normal((60, 80)))
    b = np.asfortranarray(rng.standard_normal((60, 80)))
    want = O.binary("add", a, b)
    r = ap.add(ap.asarray(a), ap.asarray(b), order=order).get()
    _same_bytes(r, want)
    # NumPy's layouts for F-ordered (60, 80) float64 operands: order C is
    # C-contiguous, F / A / K follow the operands (F-contiguous)
    want_strides = (640, 8) if order == "C" else (8, 480)
    assert r.strides == want_strides


def test_reduce_out_overlapping_input(ap):
    """sum along axis 1 written into the input's own first column."""
    x = np.random.default_rng(6).standard_normal((500, 300))
    want = x.copy()
    want[:, 0] = O.reduce("add", x, 1)
    X = ap.asarray(x)
    ap.add.reduce(X, axis=1, out=X[:, 0])
    _same_bytes(X.get(), want)


def test_float16_masked_reduce_matches_oracle(ap):
    """float16 where= reductions (the HALF loop per True run of a window),
    against the oracle, masked and unmasked."""
    x = np.random.default_rng(7).standard_normal((30, 40)).astype(np.float16)
    m = np.random.default_rng(8).random((30, 40)) < 0.5
    X = ap.asarray(x)
    _same_bytes(np.asarray(ap.add.reduce(X, axis=1, where=ap.asarray(m)).get()), O.reduce("add", x, 1, where=m))
    _same_bytes(np.asarray(ap.add.reduce(X, axis=1).get()), O.reduce("add", x, 1))


def test_stream_handle_is_the_current_stream(ap):
    """Calls launch on torch's current stream: the handle the facade passes
    is that stream's, on the default stream and inside a side stream."""
    import torch
    from anionpy_amd import _lib
    handle = lambda: _lib.stream_handle().value or 0  # noqa: E731  (the default stream is the null handle)
    assert handle() == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert handle() == s.cuda_stream == torch.cuda.current_stream().cuda_stream != 0
        x = ap.asarray(np.arange(4096, dtype=np.float64).reshape(64, 64))
        y = ap.add(x, 1.0)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    assert y.get().tobytes() == (np.arange(4096, dtype=np.float64).reshape(64, 64) + 1.0).tobytes()
