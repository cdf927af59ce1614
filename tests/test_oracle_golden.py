"""The CPU oracle against the NumPy 2.2.6 golden vectors (pins the oracle)."""
import numpy as np
import pytest

import oracle as O
from fixture_cases import args_axis, check_result, host_input, load

KINDS = ["binary", "where", "cast", "reduce", "argext", "accumulate", "reduceat", "fpe"]


def host_binary(op, a, b, dtype=None, casting="same_kind"):
    """The facade's host half of Ufunc.__call__ (dtype= loop resolution, NEP
    50 weak scalars: anionpy_amd._resolve_loop / _weak_operands, no device
    work), then the oracle's loop."""
    import anionpy_amd as ap
    code = getattr(ap._lib, "TRUE_DIVIDE" if op in ("divide", "true_divide") else op.upper())
    loop = ap._resolve_loop(code, a, b, dtype, casting)
    a, b, const = ap._weak_operands(a, b, code, loop)
    if const is not None:
        return np.full(np.shape(a), const, dtype=np.bool_)
    if loop is not None:
        a, b = (x if x.dtype == loop else O.cast(np.asarray(x), loop) for x in (a, b))
    return O.binary(op, np.asarray(a), np.asarray(b))


def run_oracle(case, npz):
    ins = [host_input(s, npz) for s in case["inputs"]]
    k, op, args = case["kind"], case["op"], case["args"]
    if k == "binary":
        return host_binary(op, ins[0], ins[1], args.get("dtype"), args.get("casting", "same_kind"))
    if k == "where":
        import anionpy_amd as ap
        x, y, _ = ap._weak_operands(ins[1], ins[2], wrap=True)
        return O.where(ins[0], np.asarray(x), np.asarray(y))
    if k == "cast":
        return O.cast(ins[0], np.dtype(args["dtype"]))
    if k == "reduce":
        x = ins[0]
        if args.get("dtype") and np.dtype(args["dtype"]).kind == "f":
            x = O.cast(x, np.dtype(args["dtype"]))  # NumPy's buffered cast, then the float loop
        r = O.reduce(op, x, args_axis(args), dtype=args.get("dtype"), initial=args.get("initial"),
                     where=ins[1] if args.get("where") else None)
        if args.get("keepdims"):
            ax = args_axis(args)
            axes = range(x.ndim) if ax is None else ((ax,) if isinstance(ax, int) else ax)
            r = np.expand_dims(r, tuple(a % x.ndim for a in axes))
        return r
    if k == "fpe":
        if op == "cast":
            return np.array([O.cast_fpe(ins[0], np.dtype(args["dtype"]))], dtype=np.int64)
        return np.array([O.binary_fpe("true_divide" if op == "divide" else op, ins[0], ins[1])], dtype=np.int64)
    if k == "reduceat":
        return O.reduceat(op, ins[0], args["indices"], axis=args["axis"], dtype=args.get("dtype"))
    if k == "accumulate":
        if op in ("cumsum", "cumprod"):
            x, ax = ins[0], args.get("axis")
            if ax is None:
                x, ax = np.ravel(x), 0  # C-order ravel (a host copy), then the scan
            return O.accumulate("add" if op == "cumsum" else "multiply", x, ax)
        return O.accumulate(op, ins[0], args["axis"], dtype=args.get("dtype"))
    r = O.argext(op == "argmax", ins[0], args.get("axis"))
    if args.get("keepdims"):
        r = np.expand_dims(r, args["axis"])
    return r


@pytest.mark.parametrize("kind", KINDS)
def test_oracle_matches_numpy_fixtures(kind):
    cases, npz = load()
    bad = []
    n = 0
    for c in cases:
        if c["kind"] != kind:
            continue
        n += 1
        exp = c["expected"]
        try:
            got = run_oracle(c, npz)
        except O.OracleError:
            if "error" not in exp:
                bad.append((c["name"], "oracle raised"))
            continue
        except (OverflowError, TypeError) as e:  # the facade's host half (weak scalars, dtype=)
            if exp.get("error") not in (type(e).__name__, "UFuncTypeError") or str(e) != exp["msg"]:
                bad.append((c["name"], f"raised {type(e).__name__}: {e}"))
            continue
        if "error" in exp:
            bad.append((c["name"], "oracle did not raise " + exp["error"]))
            continue
        # keepdims views from expand_dims carry stride 0 on the new axes
        if c["args"].get("keepdims"):
            c = dict(c, expected=dict(exp, check_strides=False))
        why = check_result(c, got, npz)
        if why:
            bad.append((c["name"], why))
    assert n > 0
    assert not bad, f"{len(bad)} of {n} mismatches, first: {bad[:5]}"


def test_window_model_matters():
    """A run longer than the 8192-element buffer: NumPy folds window partials
    in order; one pairwise tree over the whole run (the reference's
    pairwise_sum, ionp-core/src/ufunc.rs:6416) differs.  Pins that the oracle
    applies the windows and that the fixtures would catch a dropped window."""
    x = np.random.default_rng(3).uniform(-1, 1, 20000)
    windowed = O.reduce("add", x)
    whole = O.pairwise(x)
    assert windowed.tobytes() == np.add.reduce(x).tobytes()
    assert windowed != whole
    # and a sequential fold (a dropped tree) differs from both
    seq = np.float64(0)
    for v in x:
        seq = seq + v
    assert seq != windowed


def test_pairwise_closed_forms():
    """Exact small cases: sums of integers are exact in any order; a leaf of
    8 accumulators combines ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7))."""
    x = np.arange(1000, dtype=np.float64)
    assert O.pairwise(x) == 499500.0
    # bracketing-sensitive leaf: 1e16 cancels only in the documented order
    v = np.array([1.0, 1e16, -1e16, 1.0, 0.0, 0.0, 0.0, 0.0])
    r = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]))
    assert O.pairwise(v) == r
    flat = 0.0
    for t in v:
        flat += t
    assert r != flat  # a flat fold of the accumulators would be caught


def test_signed_zero_seed():
    """add.reduce seeds with +0: a sum of -0.0 is +0.0 (ufunc.rs:7137)."""
    r = O.reduce("add", np.full(5, -0.0))
    assert r == 0.0 and not np.signbit(r)


def test_argmax_first_nan_and_first_max():
    x = np.array([1.0, 5.0, 5.0, np.nan, 7.0, np.nan])
    assert O.argmax(x) == 3
    assert O.argmax(np.array([1.0, 5.0, 5.0])) == 1
    assert O.argmin(np.array([2.0, -1.0, -1.0])) == 1


def test_max_nan_rules():
    """Contiguous runs: a NaN the vector loop sees becomes the canonical
    quiet NaN; a strided run returns the first NaN's own bits."""
    neg = np.array([0xFFF8000000000456], dtype=np.uint64).view(np.float64)[0]
    x = np.arange(100, dtype=np.float64)
    x[10] = neg
    assert O.reduce("maximum", x).view(np.uint64) == 0x7FF8000000000000
    y = np.zeros(200)
    y[::2] = np.arange(100)
    y[20] = neg
    assert O.reduce("maximum", y[::2]).view(np.uint64) == 0xFFF8000000000456


def test_openmp_binary_equals_scalar_loop():
    """The CPU baseline's multi-thread leg computes the same bytes and layout
    as the scalar loop (rows split per thread; broadcast and strided views)."""
    rng = np.random.default_rng(31)
    X = rng.standard_normal((600, 500))
    cases = [(X, rng.standard_normal(500)), (X.T[::2], rng.standard_normal((250, 1))),
             (rng.integers(-2 ** 62, 2 ** 62, (97, 3)), rng.integers(-2 ** 62, 2 ** 62, (97, 1))),
             (rng.standard_normal(5).astype(np.float32), np.float64(2.0)), (np.float64(1.5), np.float64(2.0))]
    for a, b in cases:
        for op in ("add", "multiply", "less"):
            want = O.binary(op, a, b)
            got, nt = O.binary_omp(op, a, b)
            assert nt >= 1 and got.strides == want.strides
            assert got.tobytes() == want.tobytes()
