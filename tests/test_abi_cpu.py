"""C ABI checks that need no GPU: the library loads, exports every symbol
include/ionp_hip.h declares, and its host-only planning matches the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "anionpy_amd", "libionp_hip.so")


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "ionp_hip.h")).read()
    return sorted(set(re.findall(r"\b(ionp_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libionp_hip.so is not built (python -m anionpy_amd.build)")
    from anionpy_amd import _lib
    return _lib.load()


def test_every_header_symbol_is_exported(lib):
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from anionpy_amd import _lib
    assert set(header_symbols()) == set(_lib.SIGNATURES)


def _desc(a):
    from anionpy_amd import _lib
    import anionpy_amd as ap
    d = _lib.ArrayDesc()
    d.dtype = ap._CODE[a.dtype]
    d.ndim = a.ndim
    for i in range(a.ndim):
        d.shape[i] = a.shape[i]
        d.strides[i] = a.strides[i] // a.itemsize
    d.offset = 0
    d.data = 0
    return d


def test_result_type_table_matches_oracle(lib):
    import oracle as O
    out = ctypes.c_int32()
    for i, a in enumerate(O.DTYPES):
        for j, b in enumerate(O.DTYPES):
            assert lib.ionp_result_type(i, j, ctypes.byref(out)) == 0
            assert O.DTYPES[out.value] == O.result_type(a, b).type


def test_plan_binary_korder_matches_oracle(lib):
    import oracle as O
    from anionpy_amd import _lib
    rng = np.random.default_rng(0)
    base = rng.uniform(size=(6, 7, 8))
    views = [base, base.T, base[::-1, ::2], base.transpose(1, 0, 2), np.asfortranarray(base), base[:, :1, :]]
    for a in views:
        for b in (base[0, 0], base[:, 0, :1], a[::1]):
            try:
                shape = O.broadcast_shape(a.shape, b.shape)
            except O.OracleError:
                continue
            t = _lib.ArrayDesc()
            da, db = _desc(a), _desc(b)
            assert lib.ionp_plan_binary(0, ctypes.byref(da), ctypes.byref(db), ctypes.byref(t)) == 0
            assert tuple(t.shape[: t.ndim]) == shape
            got = [t.strides[i] for i in range(t.ndim) if t.shape[i] > 1]
            want = [s for s, n in zip(O.korder_strides([a, b], shape), shape) if n > 1]
            assert got == want
            np_out = np.add(a, b)
            assert got == [s // 8 for s, n in zip(np_out.strides, np_out.shape) if n > 1]


def test_errors_carry_numpy_messages(lib):
    from anionpy_amd import _lib
    t = _lib.ArrayDesc()
    a, b = _desc(np.zeros(3)), _desc(np.zeros((4, 5)))
    assert lib.ionp_plan_binary(0, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t)) == 1
    assert lib.ionp_last_error().decode() == "operands could not be broadcast together with shapes (3,) (4,5) "
    x = _desc(np.zeros((3, 4)))
    axes = (ctypes.c_int32 * 1)(2)
    assert lib.ionp_plan_reduce(0, ctypes.byref(x), axes, 1, 0, -1, ctypes.byref(t)) == 7
    assert lib.ionp_last_error().decode() == "axis 2 is out of bounds for array of dimension 2"
    axes = (ctypes.c_int32 * 2)(1, -1)
    assert lib.ionp_plan_reduce(0, ctypes.byref(x), axes, 2, 0, -1, ctypes.byref(t)) == 5
    assert lib.ionp_last_error().decode() == "duplicate value in 'axis'"
    bb = _desc(np.zeros(3, dtype=bool))
    assert lib.ionp_plan_binary(1, ctypes.byref(bb), ctypes.byref(bb), ctypes.byref(t)) == 4
    assert "boolean subtract" in lib.ionp_last_error().decode()


def test_plan_reduce_layout_and_dtype(lib):
    from anionpy_amd import _lib
    t = _lib.ArrayDesc()
    for a, axis, kd in [(np.zeros((4, 5, 6), np.int32), (0, 2), 0), (np.zeros((4, 5, 6)).T, (1,), 1),
                        (np.zeros((3, 4), np.uint8), None, 0)]:
        d = _desc(a)
        if axis is None:
            axes, n = None, -1
        else:
            axes, n = (ctypes.c_int32 * len(axis))(*axis), len(axis)
        assert lib.ionp_plan_reduce(0, ctypes.byref(d), axes, n, kd, -1, ctypes.byref(t)) == 0
        want = np.add.reduce(a, axis=axis, keepdims=bool(kd))
        assert tuple(t.shape[: t.ndim]) == want.shape
        import anionpy_amd as ap
        assert ap._NP_DTYPES[t.dtype] == want.dtype.type
        st = [t.strides[i] for i in range(t.ndim) if t.shape[i] > 1]
        assert st == [s // want.itemsize for s, n in zip(want.strides, want.shape) if n > 1]


def test_plan_reduce_where_layout_follows_mask(lib):
    """The where= mask joins the iterator's axis order: an F-ordered mask on a
    C-ordered input changes NumPy's output layout."""
    from anionpy_amd import _lib
    import anionpy_amd as ap
    t = _lib.ArrayDesc()
    x = np.zeros((4, 5, 6))
    for m in (np.ones((4, 5, 6), bool), np.asfortranarray(np.ones((4, 5, 6), bool)), np.ones(6, bool)):
        for axis in ((1,), (0, 2), (2,)):
            axes = (ctypes.c_int32 * len(axis))(*axis)
            dx, dm = _desc(x), _desc(m)
            assert lib.ionp_plan_reduce_where(0, ctypes.byref(dx), ctypes.byref(dm), axes, len(axis), 0, -1,
                                              ctypes.byref(t)) == 0
            want = np.add.reduce(x, axis=axis, where=m)
            assert tuple(t.shape[: t.ndim]) == want.shape
            assert ap._NP_DTYPES[t.dtype] == want.dtype.type
            st = [t.strides[i] for i in range(t.ndim) if t.shape[i] > 1]
            assert st == [s // 8 for s, n in zip(want.strides, want.shape) if n > 1]
    bad = _desc(np.ones((2, 6), bool))
    assert lib.ionp_plan_reduce_where(0, ctypes.byref(_desc(np.zeros(6))), ctypes.byref(bad), None, -1, 0, -1,
                                      ctypes.byref(t)) == 1


def test_reduce_where_identity_error_before_any_device_work(lib):
    from anionpy_amd import _lib
    t = _lib.ArrayDesc()
    x = _desc(np.zeros((3, 4)))
    m = _desc(np.ones((3, 4), bool))
    axes = (ctypes.c_int32 * 1)(1)
    assert lib.ionp_plan_reduce_where(4, ctypes.byref(x), ctypes.byref(m), axes, 1, 0, -1, ctypes.byref(t)) == 0
    assert lib.ionp_reduce_where(4, ctypes.byref(x), ctypes.byref(m), axes, 1, 0, None, ctypes.byref(t), None,
                                 None) == 5
    assert lib.ionp_last_error().decode() == ("reduction operation 'maximum' does not have an identity, so to use "
                                              "a where mask one has to specify 'initial'")


def test_plan_reduceat_layout_and_index_errors(lib):
    from anionpy_amd import _lib
    import anionpy_amd as ap
    t = _lib.ArrayDesc()
    base = np.zeros((4, 5, 6), np.int8)
    for a, axis in ((base, 1), (np.asfortranarray(base), 2), (base.transpose(2, 0, 1), 0), (base[:, ::2], 1)):
        idx = [0, 2, 1]
        assert lib.ionp_plan_reduceat(0, ctypes.byref(_desc(a)), axis, len(idx), -1, ctypes.byref(t)) == 0
        want = np.add.reduceat(a, idx, axis=axis)
        assert tuple(t.shape[: t.ndim]) == want.shape
        assert ap._NP_DTYPES[t.dtype] == want.dtype.type
        st = [t.strides[i] for i in range(t.ndim) if t.shape[i] > 1]
        assert st == [s // want.itemsize for s, n in zip(want.strides, want.shape) if n > 1]
    x = _desc(np.zeros(5))
    idx = (ctypes.c_int64 * 2)(8, 7)
    assert lib.ionp_reduceat(0, ctypes.byref(x), idx, 2, 0, ctypes.byref(x), None, None) == 3
    assert lib.ionp_last_error().decode() == "index 8 out-of-bounds in add.reduceat [0, 5)"
    scalar = _desc(np.zeros(()))
    assert lib.ionp_plan_reduceat(0, ctypes.byref(scalar), 0, 1, -1, ctypes.byref(t)) == 4
    assert lib.ionp_last_error().decode() == "cannot reduceat on a scalar"


def test_float16_dtype_rules(lib):
    """float16 promotion, true_divide's loop and the reduce dtype, against NumPy."""
    from anionpy_amd import _lib
    import anionpy_amd as ap
    h = ap._CODE[np.dtype(np.float16)]
    loop, out = ctypes.c_int32(), ctypes.c_int32()
    for other in ap._NP_DTYPES:
        o = ap._CODE[np.dtype(other)]
        assert lib.ionp_binary_types(3, h, o, ctypes.byref(loop), ctypes.byref(out)) == 0
        want = np.divide(np.ones(1, np.float16), np.ones(1, other)).dtype
        assert ap._NP_DTYPES[out.value] == want.type
    t = _lib.ArrayDesc()
    for op, dt in ((0, np.float16), (2, np.float16), (4, np.float16)):
        a = np.zeros((3, 4), dt)
        axes = (ctypes.c_int32 * 1)(1)
        assert lib.ionp_plan_reduce(op, ctypes.byref(_desc(a)), axes, 1, 0, -1, ctypes.byref(t)) == 0
        assert ap._NP_DTYPES[t.dtype] == np.float16


def test_can_cast_table_matches_numpy(lib):
    """ionp_can_cast against numpy.can_cast over every dtype pair and rule
    (the rules are numpy's casting table; it has not changed in NumPy 2)."""
    import anionpy_amd as ap
    from anionpy_amd import _lib
    ok = ctypes.c_int32()
    for a in ap._NP_DTYPES:
        for b in ap._NP_DTYPES:
            for rule, code in _lib.CASTING.items():
                assert lib.ionp_can_cast(ap._CODE[np.dtype(a)], ap._CODE[np.dtype(b)], code, ctypes.byref(ok)) == 0
                assert bool(ok.value) == np.can_cast(a, b, rule), (a, b, rule)


@pytest.mark.parametrize("op", ["add", "subtract", "multiply", "true_divide", "maximum", "less", "equal"])
def test_resolve_dtype_matches_numpy_loops(lib, op):
    """ionp_resolve_dtype against NumPy 2.2.6's ufunc resolution for
    dtype= on every (a, b, requested) triple of same_kind / safe / unsafe
    casting: the loop dtype (the result's dtype, or the inputs' for a
    comparison), or the error class and message."""
    import warnings
    import anionpy_amd as ap
    from anionpy_amd import _lib
    if np.__version__ != "2.2.6":
        pytest.skip("expectations are NumPy 2.2.6's")
    code = getattr(_lib, op.upper())
    f = getattr(np, op)
    loop = ctypes.c_int32()
    dts = [np.dtype(d) for d in ap._NP_DTYPES]
    for a in dts:
        for b in dts[::3] + [a]:
            for req in dts:
                for rule in ("same_kind", "safe", "unsafe"):
                    rc = lib.ionp_resolve_dtype(code, ap._CODE[a], ap._CODE[b], ap._CODE[req], _lib.CASTING[rule],
                                                ctypes.byref(loop))
                    try:
                        with warnings.catch_warnings():
                            warnings.simplefilter("ignore")
                            r = f(np.zeros(1, a), np.ones(1, b), dtype=req, casting=rule)
                    except TypeError as e:
                        assert rc != 0, (op, a, b, req, rule, str(e))
                        msg = lib.ionp_last_error().decode()
                        assert msg == str(e), (op, a, b, req, rule)
                        assert (rc == 6) == (type(e).__name__ == "UFuncTypeError" or "Cannot cast" in str(e))
                        continue
                    assert rc == 0, (op, a, b, req, rule, lib.ionp_last_error().decode())
                    assert r.dtype == req
                    if code < _lib.LESS:
                        assert np.dtype(ap._NP_DTYPES[loop.value]) == req
