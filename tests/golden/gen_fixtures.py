"""Write the golden vectors for the ufunc / reduction hot path.

Run in a container with NumPy 2.2.6 on x86-64 -- the oracle the reference's
differential suite compares against byte-for-byte
(tests/differential/harness.py:571-630 `_bit_exact_equal`):

    python tests/golden/gen_fixtures.py

Outputs (committed):
  tests/golden/fixtures.npz   input buffers and expected outputs
  tests/golden/cases.json     case list: op, arguments, operand views
                              (offset/shape/strides in elements into a
                              stored buffer), expected dtype/shape/strides
                              or expected exception
  tests/golden/big.json       SHA-256 digests / scalars of the full-size
                              workloads (inputs regenerated from their seeds)

Cases mirror the reference's own tests: the Rust KATs of ufunc.rs
(add_same_shape, add_broadcasts_row_vector_over_matrix, add_promotes_dtype,
add_on_transposed_view_matches_manual_layout, multiply_bool_is_logical_and;
ionp-core/src/ufunc.rs:10912-10986), the reduce length sweep of
tests/differential/ufunc_cases.py:156-252 (lengths 2..1000 over 1-D,
strided 1-D, (n,3) C and F, RandomState seeds as there), the corpus'
special values (tests/differential/corpus.py:100-158: NaN, +-inf, +-0,
denormals, integer extremes) and views (negative strides, transposes,
F order, offset slices), plus the buffer-window boundaries (8191/8192/8193,
2^20+13) that the reference's sweep never reaches.
"""
import ctypes
import hashlib
import json
import os
import sys
import warnings

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
assert np.__version__ == "2.2.6", f"fixtures are pinned to NumPy 2.2.6, found {np.__version__}"
assert np.getbufsize() == 8192

warnings.simplefilter("ignore")

DTYPES = [np.bool_, np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64,
          np.float32, np.float64]


class Gen:
    """A large input regenerated from a seed instead of stored:
    (shift + scale * np.random.default_rng(seed).uniform(-1, 1, n)).astype(dtype),
    then [::step] (1-D) or reshaped to `shape`."""

    def __init__(self, seed, n, dtype, step=1, shift=None, scale=None, shape=None):
        self.seed, self.n, self.dtype, self.step = seed, n, np.dtype(dtype), step
        self.shift, self.scale, self.shape = shift, scale, shape

    def array(self):
        x = np.random.default_rng(self.seed).uniform(-1, 1, self.n)
        if self.scale is not None:
            x = (self.shift or 0.0) + self.scale * x
        x = x.astype(self.dtype)
        return x.reshape(self.shape) if self.shape is not None else x[:: self.step]

    def spec(self):
        d = {"gen": "uniform", "seed": self.seed, "n": self.n, "dtype": self.dtype.name, "step": self.step}
        if self.scale is not None:
            d.update(shift=self.shift or 0.0, scale=self.scale)
        if self.shape is not None:
            d["shape"] = list(self.shape)
        return d


class PyScalar:
    """A Python bool / int / float operand (a NEP 50 weak scalar): recorded
    as its type and exact value, passed to the ufunc as the scalar itself."""

    def __init__(self, v):
        self.v = v

    def spec(self):
        v = self.v
        if isinstance(v, bool):
            return {"py": "bool", "value": str(v)}
        if isinstance(v, int):
            return {"py": "int", "value": str(v)}
        return {"py": "float", "value": float(v).hex()}


class Store:
    def __init__(self):
        self.arrays = {}
        self.cases = []
        self.roots = {}
        self.keep = []

    def buf(self, a):
        key = f"b{len(self.arrays)}"
        self.arrays[key] = np.ascontiguousarray(a).reshape(-1)
        return key

    def view(self, v):
        """Record an operand view as (stored root buffer, element offset,
        shape, strides): views of one array share its stored buffer."""
        if isinstance(v, (Gen, PyScalar)):
            return v.spec()
        v = np.asarray(v)
        isz = v.itemsize
        st = [s // isz for s in v.strides]
        root = v
        while isinstance(root.base, np.ndarray):
            root = root.base
        if root.dtype != v.dtype or not (root.flags.c_contiguous or root.flags.f_contiguous):
            root = np.ascontiguousarray(v)
            st = [s // isz for s in root.strides]
            v = root
        key = self.roots.get(id(root))
        if key is None:
            key = self.buf(root.ravel(order="K"))  # memory order
            self.roots[id(root)] = key
            self.keep.append(root)
        off = (v.__array_interface__["data"][0] - root.__array_interface__["data"][0]) // isz
        return {"buf": key, "offset": int(off), "shape": list(v.shape), "strides": st, "dtype": v.dtype.name}

    def expect(self, r):
        """Expected output: its values (C order) or, above 8 KB, their SHA-256."""
        r = np.asarray(r)
        isz = r.itemsize
        e = {"dtype": r.dtype.name, "shape": list(r.shape), "strides": [s // isz for s in r.strides]}
        if r.nbytes > 8192:
            e["sha256"] = sha(r)
        else:
            e["buf"] = self.buf(np.ascontiguousarray(r))
        return e

    def add(self, name, kind, op, inputs, fn, args=None, check_strides=True):
        try:
            r = fn()
            exp = self.expect(r)
            exp["check_strides"] = check_strides
        except Exception as e:  # noqa: BLE001 - expected exceptions are data here
            exp = {"error": type(e).__name__, "msg": str(e)}
        self.cases.append({"name": name, "kind": kind, "op": op, "args": args or {},
                           "inputs": [self.view(x) for x in inputs], "expected": exp})


def specials(dt, n, rng):
    dt = np.dtype(dt)
    if dt.kind == "f":
        base = rng.standard_normal(n) * 10
        sp = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0, 1.0, -1.0, 0.5, 1e-310, -1e-310,
                       1e300, -1e300, 3.4e38, 1e-45, 2.5, -2.5], dtype=np.float64)
        base[: min(n, len(sp))] = sp[: min(n, len(sp))]
        rng.shuffle(base)
        return base.astype(dt)
    if dt.kind == "b":
        return rng.random(n) < 0.5
    info = np.iinfo(dt)
    base = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
    sp = np.array([info.min, info.max, 0, 1, info.max - 1, info.min + 1], dtype=dt)
    base[: min(n, len(sp))] = sp[: min(n, len(sp))]
    rng.shuffle(base)
    return base


def binary_cases(S, rng):
    ops = ["add", "subtract", "multiply", "true_divide", "maximum", "minimum", "less", "less_equal",
           "greater", "greater_equal", "equal", "not_equal"]
    # reference Rust KATs (ufunc.rs:10912-10986)
    a = np.array([1.0, 2.0, 3.0]); b = np.array([10.0, 20.0, 30.0])
    S.add("kat/add_same_shape", "binary", "add", [a, b], lambda: np.add(a, b))
    a = np.arange(6.0).reshape(2, 3); b = np.array([100.0, 200.0, 300.0])
    S.add("kat/add_broadcasts_row_vector_over_matrix", "binary", "add", [a, b], lambda: np.add(a, b))
    a = np.array([1, 2, 3], dtype=np.int32); b = np.array([0.5, 0.5, 0.5])
    S.add("kat/add_promotes_dtype", "binary", "add", [a, b], lambda: np.add(a, b))
    a = np.array([True, True, False]); b = np.array([True, False, False])
    S.add("kat/multiply_bool_is_logical_and", "binary", "multiply", [a, b], lambda: np.multiply(a, b))
    a = np.arange(6.0).reshape(2, 3).T; b = np.array([1.0, 1.0])
    S.add("kat/add_on_transposed_view_matches_manual_layout", "binary", "add", [a, b], lambda: np.add(a, b))

    # every op over same-dtype pairs with special values, several layouts
    for op in ops:
        f = getattr(np, op)
        for dt in DTYPES:
            if op == "subtract" and dt == np.bool_:
                x = specials(dt, 4, rng)
                S.add(f"binary/{op}/{np.dtype(dt).name}/error", "binary", op, [x, x], lambda: f(x, x))
                continue
            A = specials(dt, 7 * 9, rng).reshape(7, 9)
            B = specials(dt, 9, rng)
            C = specials(dt, 7, rng).reshape(7, 1)
            nm = np.dtype(dt).name
            S.add(f"binary/{op}/{nm}/contig", "binary", op, [A, A[::-1]], lambda: f(A, A[::-1]))
            S.add(f"binary/{op}/{nm}/row", "binary", op, [A, B], lambda: f(A, B))
            S.add(f"binary/{op}/{nm}/col", "binary", op, [A, C], lambda: f(A, C))
            S.add(f"binary/{op}/{nm}/T", "binary", op, [A.T, A.T[:, ::-1]], lambda: f(A.T, A.T[:, ::-1]))
            S.add(f"binary/{op}/{nm}/strided", "binary", op, [A[::2, 1::3], B[::3]], lambda: f(A[::2, 1::3], B[::3]))
            S.add(f"binary/{op}/{nm}/0d", "binary", op, [A, np.array(B[2])], lambda: f(A, np.array(B[2])))
    # mixed dtypes (promotion + fused casts), including int64 vs uint64 comparisons
    pairs = [(a, b) for a in DTYPES for b in DTYPES if a != b]
    for i, (da, db) in enumerate(pairs):
        x = specials(da, 30, rng).reshape(5, 6)
        y = specials(db, 6, rng)
        for op in (["add", "multiply", "less", "maximum"] if i % 2 == 0 else ["subtract", "equal", "true_divide", "minimum", "greater_equal"]):
            f = getattr(np, op)
            S.add(f"mixed/{op}/{np.dtype(da).name}-{np.dtype(db).name}", "binary", op, [x, y], lambda: f(x, y))
    # broadcast shapes that must fail (corpus must-fail pairs)
    x = np.zeros(3); y = np.zeros((4, 5))
    S.add("error/broadcast", "binary", "add", [x, y], lambda: np.add(x, y))
    # larger layouts reaching the vectorised row path
    for dt in (np.float32, np.float64, np.int64, np.int32, np.uint8):
        A = specials(dt, 21 * 1029, rng).reshape(21, 1029)
        b = specials(dt, 1029, rng)
        c = specials(dt, 21, rng).reshape(21, 1)
        nm = np.dtype(dt).name
        for op in ("add", "multiply", "less", "maximum"):
            f = getattr(np, op)
            S.add(f"rows/{op}/{nm}/row", "binary", op, [A, b], lambda: f(A, b))
            S.add(f"rows/{op}/{nm}/col", "binary", op, [A, c], lambda: f(A, c))
        A2 = specials(dt, 16 * 2048, rng).reshape(16, 2048)
        c2 = specials(dt, 16, rng).reshape(16, 1)
        S.add(f"rows/multiply/{nm}/aligned", "binary", "multiply", [A2, c2], lambda: np.multiply(A2, c2))
        S.add(f"rows/multiply/{nm}/offset1", "binary", "multiply", [A2[:, 1:], c2], lambda: np.multiply(A2[:, 1:], c2))
        rows_view_cases(S, A2, nm, dt)
    transpose_cases(S)


def rows_view_cases(S, A2, nm, dt):
    """The transposed stepped view of the bench (X.T[::2] * c[::2]) on a
    (16, 2048) array: A2.T[::2] is (1024, 16) with element strides (2, 2048)
    (the stride-2 gather of the row kernel), against a (1024, 1) column and
    a (16,) row; an odd element offset (A2.T[1::2]: 16-byte pairs not
    aligned) and stride 3 (A2.T[::3]: element by element).  Partners from
    their own generator (seed 90 + dtype)."""
    rv = np.random.default_rng(90 + np.dtype(dt).num)
    for vn, V in (("Tview", A2.T[::2]), ("Tview_odd", A2.T[1::2]), ("Tview_s3", A2.T[::3])):
        col = specials(dt, V.shape[0], rv).reshape(V.shape[0], 1)
        row = specials(dt, V.shape[1], rv)
        S.add(f"rows/multiply/{nm}/{vn}/col", "binary", "multiply", [V, col], lambda: np.multiply(V, col))
        S.add(f"rows/add/{nm}/{vn}/row", "binary", "add", [V, row], lambda: np.add(V, row))
        S.add(f"rows/less/{nm}/{vn}/col", "binary", "less", [col, V], lambda: np.less(col, V))


def transpose_cases(S):
    """Operands whose contiguous axes differ (A + B.T, a 3-D batch of
    transposes, negative and stepped strides, a broadcast partner), with
    extents that are not multiples of 64.  Own generator (seed 79)."""
    r4 = np.random.default_rng(79)
    ops = ["add", "subtract", "multiply", "true_divide", "maximum", "minimum", "less", "less_equal", "greater",
           "greater_equal", "equal", "not_equal"]
    for dt in (np.float64, np.float32, np.int64, np.int32):
        nm = np.dtype(dt).name
        A = specials(dt, 70 * 90, r4).reshape(70, 90)
        B = specials(dt, 90 * 70, r4).reshape(90, 70)
        C = specials(dt, 3 * 66 * 130, r4).reshape(3, 66, 130)
        D = specials(dt, 3 * 130 * 66, r4).reshape(3, 130, 66)
        row = specials(dt, 90, r4).reshape(1, 90)
        for k, op in enumerate(ops):
            if op == "true_divide" and dt in (np.int64, np.int32):
                continue
            f = getattr(np, op)
            S.add(f"transpose/{op}/{nm}/A+B.T", "binary", op, [A, B.T], lambda: f(A, B.T))
            S.add(f"transpose/{op}/{nm}/B.T+A", "binary", op, [B.T, A], lambda: f(B.T, A))
            if k % 3 == 0:
                S.add(f"transpose/{op}/{nm}/batchT", "binary", op, [C, D.transpose(0, 2, 1)],
                      lambda: f(C, D.transpose(0, 2, 1)))
            if k % 3 == 1:
                S.add(f"transpose/{op}/{nm}/negT", "binary", op, [A[::-1], B.T[:, ::-1]], lambda: f(A[::-1], B.T[:, ::-1]))
            if k % 3 == 2:
                S.add(f"transpose/{op}/{nm}/rowT", "binary", op, [row, B.T], lambda: f(row, B.T))
                S.add(f"transpose/{op}/{nm}/stepT", "binary", op, [A[::2, 1::2], B.T[::2, 1::2]],
                      lambda: f(A[::2, 1::2], B.T[::2, 1::2]))


def where_cast_cases(S, rng):
    c = rng.random((4, 5)) < 0.5
    x = rng.uniform(size=(5,)).astype(np.float32)
    y = np.arange(4, dtype=np.int64)[:, None]
    S.add("where/mixed", "where", "where", [c, x, y], lambda: np.where(c, x, y))
    ci = rng.integers(0, 3, (6, 7)).astype(np.int32)
    xf = specials(np.float64, 42, rng).reshape(6, 7)
    S.add("where/intcond_T", "where", "where", [ci.T, xf.T, np.array(2.5)], lambda: np.where(ci.T, xf.T, np.array(2.5)))
    for dx in (np.float32, np.float64, np.int64, np.uint8):
        x = specials(dx, 3 * 2000, rng).reshape(3, 2000)
        y = specials(dx, 2000, rng)
        cc = rng.random((3, 2000)) < 0.3
        S.add(f"where/{np.dtype(dx).name}", "where", "where", [cc, x, y], lambda: np.where(cc, x, y))
    S.add("error/where_broadcast", "where", "where", [np.zeros(3, bool), np.zeros(4), np.zeros(3)],
          lambda: np.where(np.zeros(3, bool), np.zeros(4), np.zeros(3)))
    for s in DTYPES:
        for d in DTYPES:
            x = specials(s, 64, rng)
            if np.dtype(s).kind == "f" and np.dtype(d) == np.uint32:
                # float -> uint32 out of range takes different x86 paths in NumPy
                # (vector and scalar loops disagree): only the defined range is pinned
                x = np.abs(np.nan_to_num(x, nan=1.0, posinf=2.0, neginf=3.0)) % 4.0e9
                x = x.astype(s)
            S.add(f"cast/{np.dtype(s).name}->{np.dtype(d).name}", "cast", "astype", [x],
                  lambda: x.astype(d), args={"dtype": np.dtype(d).name})
            if s == np.float64:
                xv = x[::-3]
                S.add(f"cast/{np.dtype(s).name}->{np.dtype(d).name}/neg", "cast", "astype", [xv],
                      lambda: xv.astype(d), args={"dtype": np.dtype(d).name})


def reduce_cases(S, rng):
    # the reference's reduce length sweep (ufunc_cases.py:156-252), add only
    for dtype_i, dtype in enumerate(("float32", "float64")):
        for n in (2, 4, 7, 8, 9, 16, 32, 64, 127, 128, 129, 1000):
            seed = (0 * 10_000 + (dtype_i + 1) * 1_000 + n) % (2 ** 31)
            mk = lambda k, sd: np.random.RandomState(sd).uniform(-4.0, 4.0, size=k).astype(np.float64)  # noqa: E731
            flat = mk(n, seed).astype(dtype)
            S.add(f"sweep/add/{dtype}/n{n}/1d", "reduce", "add", [flat], lambda: np.add.reduce(flat), {"axis": 0})
            S.add(f"sweep/add/{dtype}/n{n}/1d/none", "reduce", "add", [flat], lambda: np.add.reduce(flat, axis=None),
                  {"axis": None})
            nc = mk(2 * n, seed + 2).astype(dtype)[::2]
            S.add(f"sweep/add/{dtype}/n{n}/1d_noncontig", "reduce", "add", [nc], lambda: np.add.reduce(nc), {"axis": 0})
            b2 = mk(n * 3, seed + 4).reshape(n, 3)
            c2 = np.ascontiguousarray(b2.astype(dtype))
            f2 = np.asfortranarray(b2.astype(dtype))
            S.add(f"sweep/add/{dtype}/n{n}/2d_c", "reduce", "add", [c2], lambda: np.add.reduce(c2), {"axis": 0})
            S.add(f"sweep/add/{dtype}/n{n}/2d_f", "reduce", "add", [f2], lambda: np.add.reduce(f2), {"axis": 0})
            S.add(f"sweep/maximum/{dtype}/n{n}/1d", "reduce", "maximum", [flat], lambda: np.maximum.reduce(flat),
                  {"axis": 0})
    # buffer-window boundaries
    for dt in (np.float32, np.float64):
        nm = np.dtype(dt).name
        for n in (8191, 8192, 8193, 12345, 16384, 20000, 2 ** 20 + 13):
            g = Gen(n, n, dt)
            x = g.array()
            S.add(f"window/add/{nm}/n{n}", "reduce", "add", [g], lambda: np.add.reduce(x), {"axis": 0})
            g2 = Gen(n + 1, 2 * n, dt, -2)
            xs = g2.array()
            S.add(f"window/add/{nm}/n{n}/neg2", "reduce", "add", [g2], lambda: np.add.reduce(xs), {"axis": 0})
            S.add(f"window/maximum/{nm}/n{n}/neg2", "reduce", "maximum", [g2], lambda: np.maximum.reduce(xs), {"axis": 0})
        for shp in ((2, 17000), (9000, 3), (64, 300), (5, 3000), (3, 9000)):
            A = rng.uniform(-1, 1, shp).astype(dt)
            for vn, V in (("C", A), ("F", np.asfortranarray(A)), ("T", A.T), ("sub", A[:, : shp[1] - 1]),
                          ("rev", A[::-1, ::2])):
                for ax in (0, 1, None, (0, 1)):
                    S.add(f"layout/add/{nm}/{shp}/{vn}/{ax}", "reduce", "add", [V],
                          lambda: np.add.reduce(V, axis=ax), {"axis": ax})
        for shp in ((3, 60, 160), (4, 50, 200), (6, 7, 9)):
            A = rng.uniform(-1, 1, shp).astype(dt)
            for vn, V in (("C", A), ("perm", A.transpose(2, 0, 1)), ("sub", A[:, ::2, : shp[2] - 3])):
                for ax in ((1, 2), (0, 2), (0, 1), None, 1, (0, 1, 2)):
                    for kd in (False, True):
                        if kd and ax not in ((0, 2), None):
                            continue
                        S.add(f"layout3/add/{nm}/{shp}/{vn}/{ax}/kd{int(kd)}", "reduce", "add", [V],
                              lambda: np.add.reduce(V, axis=ax, keepdims=kd), {"axis": ax, "keepdims": kd})
        # specials: NaN / inf / -0
        x = rng.uniform(-1, 1, 9000).astype(dt)
        x[17] = np.inf; x[4000] = -np.inf
        S.add(f"special/add/{nm}/inf-inf", "reduce", "add", [x], lambda: np.add.reduce(x), {"axis": 0})
        x = rng.uniform(-1, 1, 9000).astype(dt)
        x[8500] = np.nan
        S.add(f"special/add/{nm}/nan", "reduce", "add", [x], lambda: np.add.reduce(x), {"axis": 0})
        x = np.full(5, -0.0, dtype=dt)
        S.add(f"special/add/{nm}/negzero", "reduce", "add", [x], lambda: np.add.reduce(x), {"axis": 0})
        A = rng.uniform(-1, 1, (300, 5)).astype(dt)
        A[7, 1] = np.nan; A[9, 2] = np.inf; A[11, 2] = -np.inf
        S.add(f"special/add/{nm}/narrow_nan", "reduce", "add", [A], lambda: np.add.reduce(A, axis=0), {"axis": 0})
        # float sums with a wider dtype (buffered casts)
        S.add(f"dtype/add/{nm}->float64", "reduce", "add", [A], lambda: np.add.reduce(A, axis=1, dtype=np.float64),
              {"axis": 1, "dtype": "float64"})

    # maximum / minimum with +-0 ties and NaN payloads (x86 AVX-512 loop)
    def f64(u):
        return np.array([u], dtype=np.uint64).view(np.float64)[0]

    P = f64(0x7ff8000000000123); N = f64(0xfff8000000000456); Sn = f64(0x7ff4000000000789)
    for trial in range(160):
        dt = (np.float64, np.float32)[trial % 2]
        nd = 1 + trial % 3
        shp = tuple(int(v) for v in rng.integers(1, [9000 if trial % 8 else 17000, 100, 24][nd - 1], nd))
        x = np.zeros(shp, dtype=dt)
        m = rng.random(shp)
        x[m < 0.5] = -0.0
        x[m < 0.2] = -1
        if trial % 4 == 1:
            x[rng.random(shp) < 0.01] = [P, N, Sn][trial % 3]
        if trial % 4 == 3:
            x = rng.uniform(-1, 1, shp).astype(dt)
            x[rng.random(shp) < 0.001] = np.nan
        v = [x, x[..., ::2], x[::-1], x.T if nd > 1 else x[::3]][trial % 4 if x.shape[-1] > 1 else 0]
        ax = [None] + list(range(v.ndim))
        ax = ax[trial % len(ax)]
        op = ("maximum", "minimum")[(trial // 2) % 2]
        S.add(f"ext/{op}/{np.dtype(dt).name}/{trial}", "reduce", op, [v],
              lambda: getattr(np, op).reduce(v, axis=ax), {"axis": ax})
    S.add("error/max_empty", "reduce", "maximum", [np.zeros((3, 0))], lambda: np.maximum.reduce(np.zeros((3, 0)), axis=1),
          {"axis": 1})
    S.add("ext/max_empty_kept", "reduce", "maximum", [np.zeros((0, 3))], lambda: np.maximum.reduce(np.zeros((0, 3)), axis=1),
          {"axis": 1})
    S.add("reduce/add_empty", "reduce", "add", [np.zeros((3, 0))], lambda: np.add.reduce(np.zeros((3, 0)), axis=1),
          {"axis": 1})
    S.add("error/axis", "reduce", "add", [np.zeros((3, 4))], lambda: np.add.reduce(np.zeros((3, 4)), axis=2), {"axis": 2})
    S.add("error/dup_axis", "reduce", "add", [np.zeros((3, 4))], lambda: np.add.reduce(np.zeros((3, 4)), axis=(1, -1)),
          {"axis": [1, -1]})
    # integer and bool reductions (reduce dtype widening)
    for dt in (np.int32, np.int64, np.uint8, np.int8, np.uint32, np.bool_, np.int16, np.uint64):
        x = specials(dt, 50 * 70, rng).reshape(50, 70)
        nm = np.dtype(dt).name
        for ax in (0, 1, None):
            for op in ("add", "maximum", "minimum"):
                S.add(f"int/{op}/{nm}/{ax}", "reduce", op, [x], lambda: getattr(np, op).reduce(x, axis=ax), {"axis": ax})
        S.add(f"int/add/{nm}/T", "reduce", "add", [x.T[::2]], lambda: np.add.reduce(x.T[::2], axis=1), {"axis": 1})
    x = rng.integers(-1000, 1000, 150000, dtype=np.int32)
    S.add("int/add/int32/long", "reduce", "add", [x], lambda: np.add.reduce(x), {"axis": 0})
    # explicit narrow result dtypes: add on bool is logical or (ufunc.rs:2341),
    # integer results wrap in the requested width (own generator: the cases
    # above keep their draws)
    r2 = np.random.default_rng(77)
    b = r2.random((300, 700)) < 0.6
    b[:, 5] = False
    for ax in (0, 1, None):
        S.add(f"int/add/bool->bool/{ax}", "reduce", "add", [b], lambda: np.add.reduce(b, axis=ax, dtype=np.bool_),
              {"axis": ax, "dtype": "bool"})
    x = r2.integers(-2 ** 31, 2 ** 31, (40, 900), dtype=np.int32)
    for ax in (0, 1):
        S.add(f"int/add/int32->int8/{ax}", "reduce", "add", [x], lambda: np.add.reduce(x, axis=ax, dtype=np.int8),
              {"axis": ax, "dtype": "int8"})
    u = r2.integers(0, 256, (90, 333), dtype=np.uint8)
    S.add("int/add/uint8->int16/1", "reduce", "add", [u], lambda: np.add.reduce(u, axis=1, dtype=np.int16),
          {"axis": 1, "dtype": "int16"})
    i8 = r2.integers(-5, 5, (4, 6), dtype=np.int8)
    S.add("int/add/int8->bool/1", "reduce", "add", [i8], lambda: np.add.reduce(i8, axis=1, dtype=np.bool_),
          {"axis": 1, "dtype": "bool"})
    for op in ("maximum", "minimum"):
        S.add(f"int/{op}/int32->int8/1", "reduce", op, [x], lambda: getattr(np, op).reduce(x, axis=1, dtype=np.int8),
              {"axis": 1, "dtype": "int8"})
        S.add(f"int/{op}/uint8->int8/1", "reduce", op, [u], lambda: getattr(np, op).reduce(u, axis=1, dtype=np.int8),
              {"axis": 1, "dtype": "int8"})
    initial_cases(S)


def initial_cases(S):
    """reduce with initial= (reduce_axis's `initial`, ufunc.rs:7634): add
    starts the window fold from it; maximum/minimum feed every element to the
    loop; an empty reduction returns it.  Own generator (seed 78)."""
    r3 = np.random.default_rng(78)

    def red(op, x, ax, init, **kw):
        return getattr(np, op).reduce(x, axis=ax, initial=init, **kw)

    for dt in (np.float64, np.float32):
        nm = np.dtype(dt).name
        x = r3.uniform(-1, 1, 20000).astype(dt)
        S.add(f"init/add/{nm}/wide", "reduce", "add", [x], lambda: red("add", x, 0, 1e3), {"axis": 0, "initial": 1e3})
        g = Gen(78, 2 ** 20 + 13, dt)
        xg = g.array()
        S.add(f"init/add/{nm}/2^20+13", "reduce", "add", [g], lambda: red("add", xg, 0, 7.0),
              {"axis": 0, "initial": 7.0})
        A = r3.uniform(-1, 1, (3000, 4)).astype(dt) * 1e4
        S.add(f"init/add/{nm}/narrow", "reduce", "add", [A], lambda: red("add", A, 0, 0.25), {"axis": 0, "initial": 0.25})
        B = r3.uniform(-1, 1, (500, 50)).astype(dt)
        S.add(f"init/add/{nm}/small", "reduce", "add", [B], lambda: red("add", B, 1, -2.5), {"axis": 1, "initial": -2.5})
        S.add(f"init/add/{nm}/all", "reduce", "add", [B], lambda: red("add", B, None, 1.5), {"axis": None, "initial": 1.5})
        for n in (1, 5, 20):
            z = np.full(n, -0.0, dtype=dt)
            S.add(f"init/add/{nm}/negzero{n}", "reduce", "add", [z], lambda: red("add", z, 0, -0.0),
                  {"axis": 0, "initial": -0.0})
        for op in ("maximum", "minimum"):
            y = r3.uniform(-1, 1, 17000).astype(dt)
            S.add(f"init/{op}/{nm}/contig", "reduce", op, [y], lambda: red(op, y, 0, 0.5), {"axis": 0, "initial": 0.5})
            S.add(f"init/{op}/{nm}/strided", "reduce", op, [y[::3]], lambda: red(op, y[::3], 0, -0.5),
                  {"axis": 0, "initial": -0.5})
            yn = y.copy()
            yn[r3.random(17000) < 0.001] = np.nan
            S.add(f"init/{op}/{nm}/nan", "reduce", op, [yn], lambda: red(op, yn, 0, 0.5), {"axis": 0, "initial": 0.5})
            N = r3.uniform(-1, 1, (4000, 3)).astype(dt)
            S.add(f"init/{op}/{nm}/narrow", "reduce", op, [N], lambda: red(op, N, 0, 0.9), {"axis": 0, "initial": 0.9})
            zz = np.where(r3.random(100) < 0.5, -0.0, 0.0).astype(dt)
            for iv in (0.0, -0.0):
                S.add(f"init/{op}/{nm}/zeros/{iv}", "reduce", op, [zz], lambda: red(op, zz, 0, iv),
                      {"axis": 0, "initial": iv})
            E = np.zeros((0, 3), dtype=dt)
            S.add(f"init/{op}/{nm}/empty", "reduce", op, [E], lambda: red(op, E, 0, 1.5), {"axis": 0, "initial": 1.5})
    E = np.zeros((3, 0))
    S.add("init/add/empty", "reduce", "add", [E], lambda: red("add", E, 1, 2.0), {"axis": 1, "initial": 2.0})
    I = r3.integers(-1000, 1000, (40, 900), dtype=np.int32)
    S.add("init/add/int32", "reduce", "add", [I], lambda: red("add", I, 1, 5), {"axis": 1, "initial": 5})
    S.add("init/maximum/int32", "reduce", "maximum", [I], lambda: red("maximum", I, 0, 990), {"axis": 0, "initial": 990})
    S.add("init/minimum/int32/all", "reduce", "minimum", [I], lambda: red("minimum", I, None, -2000),
          {"axis": None, "initial": -2000})
    U = r3.integers(0, 256, (30, 200), dtype=np.uint8)
    S.add("init/add/uint8->uint8", "reduce", "add", [U], lambda: red("add", U, 1, 250, dtype=np.uint8),
          {"axis": 1, "initial": 250, "dtype": "uint8"})
    Bo = r3.random((20, 30)) < 0.05
    S.add("init/add/bool->bool", "reduce", "add", [Bo], lambda: red("add", Bo, 1, True, dtype=np.bool_),
          {"axis": 1, "initial": True, "dtype": "bool"})


def accumulate_cases(S):
    """ufunc.accumulate (accumulate_axis, ufunc.rs:7967) and cumsum: the
    sequential scan per fiber, reduce-family dtypes, K-order outputs.  Own
    generator (seed 80)."""
    r5 = np.random.default_rng(80)
    for dt in (np.float64, np.float32, np.int64, np.int32, np.int8, np.uint8, np.bool_):
        nm = np.dtype(dt).name
        for shp in ((300,), (37, 41), (5, 20, 33)):
            x = specials(dt, int(np.prod(shp)), r5).reshape(shp)
            views = [("C", x), ("rev", x[::-1])]
            if len(shp) > 1:
                views += [("T", x.T), ("step", x[:, ::2])]
            for vn, v in views:
                for ax in range(v.ndim):
                    for op in ("add", "maximum", "minimum"):
                        S.add(f"acc/{op}/{nm}/{shp}/{vn}/{ax}", "accumulate", op, [v],
                              lambda: getattr(np, op).accumulate(v, axis=ax), {"axis": ax})
    # long fibers (few outputs) and many short ones, at sizes past one block
    for dt in (np.float64, np.float32):
        nm = np.dtype(dt).name
        g = Gen(80, 20000, dt)
        xg = g.array()
        for op in ("add", "maximum"):
            S.add(f"acc/{op}/{nm}/long1d", "accumulate", op, [g], lambda: getattr(np, op).accumulate(xg),
                  {"axis": 0})
        A = r5.uniform(-1, 1, (9000, 3)).astype(dt)
        A[r5.random(A.shape) < 1e-3] = np.nan
        for ax in (0, 1):
            S.add(f"acc/add/{nm}/9000x3/{ax}", "accumulate", "add", [A], lambda: np.add.accumulate(A, axis=ax),
                  {"axis": ax})
            S.add(f"acc/minimum/{nm}/9000x3/{ax}", "accumulate", "minimum", [A],
                  lambda: np.minimum.accumulate(A, axis=ax), {"axis": ax})
        gB = Gen(8000 + len(nm), 300 * 700, dt, shape=(300, 700))
        B = gB.array()
        for ax in (0, 1):
            S.add(f"acc/add/{nm}/300x700/{ax}", "accumulate", "add", [gB], lambda: np.add.accumulate(B, axis=ax),
                  {"axis": ax})
        T = r5.uniform(-1, 1, (600, 150)).astype(dt)
        T[r5.random(T.shape) < 2e-4] = np.nan
        for op in ("add", "maximum"):
            S.add(f"acc/{op}/{nm}/600x150/1", "accumulate", op, [T], lambda: getattr(np, op).accumulate(T, axis=1),
                  {"axis": 1})
    W = r5.integers(-2 ** 62, 2 ** 62, (400, 130))
    S.add("acc/add/int64/400x130/1", "accumulate", "add", [W], lambda: np.add.accumulate(W, axis=1), {"axis": 1})
    I = r5.integers(-2 ** 62, 2 ** 62, (2000, 5))
    S.add("acc/add/int64/wrap", "accumulate", "add", [I], lambda: np.add.accumulate(I, axis=0), {"axis": 0})
    J = r5.integers(-1000, 1000, (40, 50), dtype=np.int32)
    S.add("acc/add/int32->float64", "accumulate", "add", [J], lambda: np.add.accumulate(J, axis=1, dtype=np.float64),
          {"axis": 1, "dtype": "float64"})
    S.add("acc/add/int32->int16", "accumulate", "add", [J], lambda: np.add.accumulate(J, axis=0, dtype=np.int16),
          {"axis": 0, "dtype": "int16"})
    S.add("error/acc_axis", "accumulate", "add", [J], lambda: np.add.accumulate(J, axis=2), {"axis": 2})
    E = np.zeros((0, 4))
    S.add("acc/add/empty", "accumulate", "add", [E], lambda: np.add.accumulate(E, axis=0), {"axis": 0})
    # cumsum: add.accumulate over the C-order ravel when axis is None
    K = r5.standard_normal((30, 40))
    S.add("cumsum/float64/None", "accumulate", "cumsum", [K.T], lambda: np.cumsum(K.T), {"axis": None})
    S.add("cumsum/float64/1", "accumulate", "cumsum", [K], lambda: np.cumsum(K, axis=1), {"axis": 1})
    S.add("cumsum/int32/None", "accumulate", "cumsum", [J], lambda: np.cumsum(J), {"axis": None})


def prod_cases(S):
    """multiply.reduce / prod (sequential_group_ordered, ufunc.rs:6583: one
    running product in iterator order, identity 1) and multiply.accumulate /
    cumprod.  Values near 1 keep products finite.  Own generator (seed 81)."""
    r6 = np.random.default_rng(81)

    def near1(shape, dt):
        # NaN and -0.0, but no inf: inf * 0 would make a second NaN payload,
        # and which of two different NaNs NumPy keeps depends on its SIMD
        # loop variant (an unpinned detail)
        x = (1 + r6.standard_normal(shape) * 0.05).astype(dt)
        m = r6.random(shape)
        x[m < 0.01] = np.nan
        x[(m >= 0.01) & (m < 0.02)] = -0.0
        return x

    for dt in (np.float64, np.float32):
        nm = np.dtype(dt).name
        for shp in ((300,), (37, 41), (5, 20, 33)):
            x = near1(shp, dt)
            views = [("C", x), ("rev", x[::-1])] + ([("T", x.T), ("step", x[:, ::2])] if len(shp) > 1 else [])
            for vn, v in views:
                axes = [None] + list(range(v.ndim)) + ([(0, 2)] if v.ndim == 3 else [])
                for ax in axes:
                    S.add(f"prod/{nm}/{shp}/{vn}/{ax}", "reduce", "multiply", [v], lambda: np.multiply.reduce(v, axis=ax),
                          {"axis": list(ax) if isinstance(ax, tuple) else ax})
                S.add(f"cumprod/{nm}/{shp}/{vn}", "accumulate", "multiply", [v],
                      lambda: np.multiply.accumulate(v, axis=v.ndim - 1), {"axis": v.ndim - 1})
        gb = Gen(8101, 20000, dt, shift=1.0, scale=1e-4)
        big = gb.array()
        S.add(f"prod/{nm}/long1d", "reduce", "multiply", [gb], lambda: np.multiply.reduce(big), {"axis": 0})
        for k, (shp, ax) in enumerate((((9000, 3), 0), ((300, 700), 1), ((2000, 150), 1), ((300, 700), 0))):
            gA = Gen(8110 + k, int(np.prod(shp)), dt, shift=1.0, scale=1e-3, shape=shp)
            A = gA.array()
            S.add(f"prod/{nm}/{shp}/{ax}", "reduce", "multiply", [gA], lambda: np.multiply.reduce(A, axis=ax),
                  {"axis": ax})
            S.add(f"cumprod/{nm}/{shp}/{ax}", "accumulate", "multiply", [gA],
                  lambda: np.multiply.accumulate(A, axis=ax), {"axis": ax})
        o = np.full(50, 1e30, dtype=dt)
        S.add(f"prod/{nm}/overflow", "reduce", "multiply", [o], lambda: np.multiply.reduce(o), {"axis": 0})
        S.add(f"prod/{nm}/initial", "reduce", "multiply", [x], lambda: np.multiply.reduce(x, axis=1, initial=0.5),
              {"axis": 1, "initial": 0.5})
    for dt in (np.int64, np.int32, np.int8, np.uint8, np.bool_):
        nm = np.dtype(dt).name
        x = specials(dt, 30 * 40, r6).reshape(30, 40)
        for ax in (None, 0, 1):
            S.add(f"prod/{nm}/{ax}", "reduce", "multiply", [x], lambda: np.multiply.reduce(x, axis=ax), {"axis": ax})
        S.add(f"prod/{nm}/T", "reduce", "multiply", [x.T[::2]], lambda: np.multiply.reduce(x.T[::2], axis=1), {"axis": 1})
        S.add(f"cumprod/{nm}/1", "accumulate", "multiply", [x], lambda: np.multiply.accumulate(x, axis=1), {"axis": 1})
    E = np.zeros((0, 3))
    S.add("prod/empty", "reduce", "multiply", [E], lambda: np.multiply.reduce(E, axis=0), {"axis": 0})
    K = (1 + r6.standard_normal((20, 30)) * 0.01)
    S.add("cumprod/float64/None", "accumulate", "cumprod", [K.T], lambda: np.cumprod(K.T), {"axis": None})


def where_reduce_cases(S):
    """reduce with a where= mask (reduce_axis_masked, ufunc.rs:7172, as
    NumPy 2.2.6 evaluates it): per bufsize window, the reduce loop runs once
    per maximal run of True mask values; the mask joins the iterator's axis
    order.  Runs crossing window boundaries, broadcast and F-ordered masks,
    gapped axes, +-0 ties and NaN for maximum/minimum, integer wrap, empty
    groups and the no-identity error.  Own generator (seed 82)."""
    r7 = np.random.default_rng(82)

    def red(op, x, ax, m, init=None, **kw):
        if init is not None:
            kw["initial"] = init
        return getattr(np, op).reduce(x, axis=ax, where=m, **kw)

    def args(ax, init=None, **kw):
        a = {"axis": list(ax) if isinstance(ax, tuple) else ax, "where": True}
        if init is not None:
            a["initial"] = init
        a.update(kw)
        return a

    def fvals(shape, dt, op):
        if op == "multiply":  # near 1, one NaN payload only (see prod_cases)
            x = (1 + r7.standard_normal(shape) * 0.02).astype(dt)
            x[r7.random(shape) < 0.002] = np.nan
            return x
        x = r7.standard_normal(shape).astype(dt)
        u = r7.random(shape)
        x[u < 0.04] = 0.0
        x[(u >= 0.04) & (u < 0.08)] = -0.0
        x[(u >= 0.08) & (u < 0.085)] = np.nan
        return x

    fops = [("add", None), ("multiply", None), ("maximum", 0.25), ("minimum", -0.25)]
    for dt in (np.float64, np.float32):
        nm = np.dtype(dt).name
        for op, init in fops:
            x = fvals(20000, dt, op)
            for p in (0.0, 0.3, 0.97, 1.0):
                m = r7.random(20000) < p
                S.add(f"wred/{op}/{nm}/1d/{p}", "reduce", op, [x, m], lambda: red(op, x, 0, m, init), args(0, init))
            m = r7.random(20000) < 0.9
            S.add(f"wred/{op}/{nm}/1d/rev", "reduce", op, [x[::-1], m], lambda: red(op, x[::-1], 0, m, init),
                  args(0, init))
            S.add(f"wred/{op}/{nm}/1d/step", "reduce", op, [x[::2], m[:10000]],
                  lambda: red(op, x[::2], 0, m[:10000], init), args(0, init))
            X = fvals((120, 200), dt, op)
            Xf = np.asfortranarray(X)
            mc = r7.random((120, 200)) < 0.8
            mf = np.asfortranarray(r7.random((120, 200)) < 0.8)
            mrow = r7.random(200) < 0.7
            for ax in (None, 0, 1):
                S.add(f"wred/{op}/{nm}/C/{ax}", "reduce", op, [X, mc], lambda: red(op, X, ax, mc, init), args(ax, init))
                S.add(f"wred/{op}/{nm}/xF-mC/{ax}", "reduce", op, [Xf, mc], lambda: red(op, Xf, ax, mc, init),
                      args(ax, init))
                S.add(f"wred/{op}/{nm}/xC-mF/{ax}", "reduce", op, [X, mf], lambda: red(op, X, ax, mf, init),
                      args(ax, init))
                S.add(f"wred/{op}/{nm}/row/{ax}", "reduce", op, [X, mrow], lambda: red(op, X, ax, mrow, init),
                      args(ax, init))
            S.add(f"wred/{op}/{nm}/step/1", "reduce", op, [X[:, ::2], mc[:, :100]],
                  lambda: red(op, X[:, ::2], 1, mc[:, :100], init), args(1, init))
            S.add(f"wred/{op}/{nm}/keep", "reduce", op, [X, mrow], lambda: red(op, X, 1, mrow, init, keepdims=True),
                  args(1, init, keepdims=True))
            if op in ("add", "maximum"):  # rows longer than one window
                W = fvals((6, 9000), dt, op)
                mw = r7.random((6, 9000)) < 0.95
                S.add(f"wred/{op}/{nm}/wide/1", "reduce", op, [W, mw], lambda: red(op, W, 1, mw, init), args(1, init))
                S.add(f"wred/{op}/{nm}/wide/None", "reduce", op, [W, mw], lambda: red(op, W, None, mw, init),
                      args(None, init))
            G = fvals((5, 7, 300), dt, op)
            mg = r7.random((5, 7, 300)) < 0.9
            S.add(f"wred/{op}/{nm}/gapped", "reduce", op, [G, mg], lambda: red(op, G, (0, 2), mg, init),
                  args((0, 2), init))
        # +-0 ties and NaN payloads in short and long runs
        for op in ("maximum", "minimum"):
            for k in range(4):
                z = r7.choice(np.array([0.0, -0.0, 1.0, -1.0, np.nan, 2.0], dtype=dt), 3000)
                if k == 3:
                    z = r7.choice(np.array([0.0, -0.0], dtype=dt), 3000)
                mz = r7.random(3000) < (0.5, 0.9, 0.99, 0.95)[k]
                for iv in (0.0, -0.0, float("nan")) if k == 3 else (0.5,):
                    S.add(f"wred/{op}/{nm}/ties/{k}/{iv}", "reduce", op, [z, mz], lambda: red(op, z, 0, mz, iv),
                          args(0, iv))
        x = fvals(100, dt, "add")
        mi = r7.random(100) < 0.5
        S.add(f"wred/add/{nm}/initial", "reduce", "add", [x, mi], lambda: red("add", x, 0, mi, 3.0), args(0, 3.0))
        E = np.zeros((0, 4), dtype=dt)
        me = np.ones(4, dtype=bool)
        S.add(f"wred/add/{nm}/empty", "reduce", "add", [E, me], lambda: red("add", E, 0, me), args(0))
        S.add(f"wred/multiply/{nm}/empty", "reduce", "multiply", [E, me], lambda: red("multiply", E, 0, me), args(0))
        S.add(f"wred/maximum/{nm}/empty", "reduce", "maximum", [E, me], lambda: red("maximum", E, 0, me, 1.5),
              args(0, 1.5))
        S.add(f"wred/maximum/{nm}/noinit", "reduce", "maximum", [x, mi], lambda: red("maximum", x, 0, mi),
              args(0))
    for dt in (np.int64, np.int32, np.uint8, np.int8, np.bool_):
        nm = np.dtype(dt).name
        x = specials(dt, 30 * 40, r7).reshape(30, 40)
        m = r7.random((30, 40)) < 0.6
        for op in ("add", "multiply", "maximum", "minimum"):
            init = 1 if op in ("maximum", "minimum") else None
            for ax in (None, 0, 1):
                S.add(f"wred/{op}/{nm}/{ax}", "reduce", op, [x, m], lambda: red(op, x, ax, m, init), args(ax, init))
            S.add(f"wred/{op}/{nm}/T", "reduce", op, [x.T, m.T], lambda: red(op, x.T, 1, m.T, init), args(1, init))
        big = specials(dt, 70000, r7)
        mb = r7.random(70000) < 0.9
        S.add(f"wred/add/{nm}/big", "reduce", "add", [big, mb], lambda: red("add", big, 0, mb), args(0))
        S.add(f"wred/add/{nm}/int8dt", "reduce", "add", [x, m], lambda: red("add", x, 1, m, dtype=np.int8),
              args(1, dtype="int8"))


def reduceat_cases(S):
    """ufunc.reduceat (reduceat_binary_axis, ufunc.rs:8298-8460, as NumPy
    2.2.6 evaluates it): per segment the reduce loop seeded with its first
    element, one call over the rest (float add: x[i] + pairwise(rest), no
    buffer windows -- segments longer than 8192 and than 2 * 8192 included);
    non-increasing index pairs pass x[i] through; K-order outputs; index and
    axis errors.  Float add/multiply inputs carry one NaN payload only (the
    choice between two differs between NumPy's loop variants).  Own
    generator (seed 83)."""
    r8 = np.random.default_rng(83)

    def rat(op, x, idx, ax=0):
        return getattr(np, op).reduceat(x, idx, axis=ax)

    def args(idx, ax=0):
        return {"indices": [int(i) for i in idx], "axis": ax}

    def fvals(shape, dt, op):
        if op == "multiply":
            x = (1 + r8.standard_normal(shape) * 0.02).astype(dt)
            x[r8.random(shape) < 0.002] = np.nan
            return x
        x = r8.standard_normal(shape).astype(dt)
        u = r8.random(shape)
        x[u < 0.04] = 0.0
        x[(u >= 0.04) & (u < 0.08)] = -0.0
        x[(u >= 0.08) & (u < 0.083)] = np.nan
        if op in ("maximum", "minimum"):  # a second NaN payload
            y = x.view(np.uint64 if np.dtype(dt).itemsize == 8 else np.uint32)
            y[(u >= 0.08) & (u < 0.0815)] |= 5
        return x

    ops = ("add", "multiply", "maximum", "minimum")
    for dt in (np.float64, np.float32):
        nm = np.dtype(dt).name
        for op in ops:
            x = fvals(3000, dt, op)
            for name, idx in (("sorted", [0, 3, 10, 137, 900, 2999]), ("unsorted", [500, 20, 21, 21, 2000, 7]),
                              ("one", [0]), ("last", [2999]), ("tail", [1, 2, 2990])):
                S.add(f"rat/{op}/{nm}/{name}", "reduceat", op, [x], lambda: rat(op, x, idx), args(idx))
            S.add(f"rat/{op}/{nm}/step", "reduceat", op, [x[::3]], lambda: rat(op, x[::3], [0, 40, 41, 600]),
                  args([0, 40, 41, 600]))
            S.add(f"rat/{op}/{nm}/rev", "reduceat", op, [x[::-1]], lambda: rat(op, x[::-1], [0, 1000, 2500]),
                  args([0, 1000, 2500]))
            X = fvals((60, 90), dt, op)
            for ax, idx in ((0, [0, 5, 30, 31, 10]), (1, [3, 50, 89]), (-1, [0, 1]), (-2, [59])):
                S.add(f"rat/{op}/{nm}/2d/{ax}", "reduceat", op, [X], lambda: rat(op, X, idx, ax), args(idx, ax))
                Xf = np.asfortranarray(X)
                S.add(f"rat/{op}/{nm}/2dF/{ax}", "reduceat", op, [Xf], lambda: rat(op, Xf, idx, ax), args(idx, ax))
            Y = fvals((4, 50, 6), dt, op)
            S.add(f"rat/{op}/{nm}/3d", "reduceat", op, [Y], lambda: rat(op, Y, [0, 3, 30], 1), args([0, 3, 30], 1))
            S.add(f"rat/{op}/{nm}/3dT", "reduceat", op, [Y.transpose(2, 0, 1)],
                  lambda: rat(op, Y.transpose(2, 0, 1), [1, 2, 40], 2), args([1, 2, 40], 2))
        # segments longer than one and than two pairwise blocks of 8192
        g = Gen(8301, 60000, dt)
        big = g.array()
        for name, idx in (("long", [0, 5, 9000, 27000]), ("whole", [0]), ("mid", [100, 17000, 16999, 33000])):
            S.add(f"rat/add/{nm}/{name}", "reduceat", "add", [g], lambda: rat("add", big, idx), args(idx))
            S.add(f"rat/maximum/{nm}/{name}", "reduceat", "maximum", [g], lambda: rat("maximum", big, idx), args(idx))
    for dt in (np.int64, np.int32, np.uint8, np.int8, np.bool_):
        nm = np.dtype(dt).name
        x = specials(dt, 40 * 30, r8).reshape(40, 30)
        for op in ops:
            for ax, idx in ((0, [0, 7, 7, 39, 2]), (1, [5, 6, 29]), (1, [0])):
                S.add(f"rat/{op}/{nm}/{ax}/{len(idx)}", "reduceat", op, [x], lambda: rat(op, x, idx, ax),
                      args(idx, ax))
    x = np.arange(10.0)
    for name, idx, ax in (("oob", [3, 10], 0), ("neg", [-1], 0), ("order", [8, 11, 12], 0), ("axis", [0], 1),
                          ("empty", [], 0)):
        S.add(f"rat/err/{name}", "reduceat", "add", [x], lambda: rat("add", x, idx, ax), args(idx, ax))
    E = np.zeros((0, 3))
    S.add("rat/add/emptyfib", "reduceat", "add", [E], lambda: rat("add", E, [0, 2], 1), args([0, 2], 1))


def fpe_cases(S):
    """IEEE exception flags (np.errstate; ionp-py/src/lib.rs:9077 and
    ionp-core/src/fpe.rs:173-360 on the reference side), as NumPy 2.2.6
    raises them: 1 divide, 2 over, 4 under, 8 invalid, OR-ed over one call.
    Rows of the special-value grid, products and quotients at the underflow
    boundary (x86 tininess after rounding), casts to every integer width
    and float64 -> float32, planted flags in large broadcast and strided
    calls, integer true_divide.  Own generator (seed 84)."""
    r9 = np.random.default_rng(84)

    def flags_of(fn):
        got = []
        with np.errstate(all="call", call=lambda t, f: got.append(f)):
            fn()
        r = 0
        for f in got:
            r |= f
        return np.array([r], dtype=np.int64)

    ops = (("add", np.add), ("subtract", np.subtract), ("multiply", np.multiply), ("divide", np.divide))
    for dt in (np.float64, np.float32):
        nm = np.dtype(dt).name
        fi = np.finfo(dt)
        sp = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, fi.max, -fi.max, fi.tiny, -fi.tiny,
                       fi.smallest_subnormal, 2.0, 0.5, 3.0, fi.tiny * 3], dtype=dt)
        for name, fn in ops:
            for i in range(len(sp)):
                a = sp[i:i + 1]
                S.add(f"fpe/{name}/{nm}/row{i}", "fpe", name, [a, sp], lambda: flags_of(lambda: fn(a, sp)))
        for k in range(40):  # products and quotients at the underflow boundary, one element each
            u = dt(r9.uniform(1, 2))
            v = dt(fi.tiny / u * (1 + r9.uniform(-4, 4) * fi.eps))
            A, B = np.array([u]), np.array([v])
            S.add(f"fpe/multiply/{nm}/tiny{k}", "fpe", "multiply", [A, B], lambda: flags_of(lambda: np.multiply(A, B)))
            C = np.array([dt(fi.tiny * (1 + r9.uniform(-4, 4) * fi.eps) * u)])
            S.add(f"fpe/divide/{nm}/tiny{k}", "fpe", "divide", [C, A], lambda: flags_of(lambda: np.divide(C, A)))
        # large calls: rows, gather and generic layouts with one planted flag
        X = r9.standard_normal((100, 256)).astype(dt)
        c = r9.uniform(0.5, 2, (100, 1)).astype(dt)
        S.add(f"fpe/multiply/{nm}/rows/none", "fpe", "multiply", [X, c], lambda: flags_of(lambda: np.multiply(X, c)))
        Xo = X.copy()
        Xo[73, 77] = fi.max
        c2 = c.copy()
        c2[73, 0] = 4.0
        S.add(f"fpe/multiply/{nm}/rows/over", "fpe", "multiply", [Xo, c2], lambda: flags_of(lambda: np.multiply(Xo, c2)))
        Xu = X.copy()
        Xu[7, 255] = fi.tiny
        S.add(f"fpe/multiply/{nm}/rows/under", "fpe", "multiply", [Xu, c], lambda: flags_of(lambda: np.multiply(Xu, c)))
        S.add(f"fpe/multiply/{nm}/gather/over", "fpe", "multiply", [Xo.T[::2], c2.T[:, :128]],
              lambda: flags_of(lambda: np.multiply(Xo.T[::2], c2.T[:, :128])))
        Xz = X.copy()
        Xz[90, 3] = 0.0
        S.add(f"fpe/divide/{nm}/rows/div", "fpe", "divide", [c, Xz], lambda: flags_of(lambda: np.divide(c, Xz)))
        S.add(f"fpe/divide/{nm}/generic/inv", "fpe", "divide", [Xz.T, Xz[:, :1].T * 0],
              lambda: flags_of(lambda: np.divide(Xz.T, Xz[:, :1].T * 0)))
        Xi = X.copy()
        Xi[0, 0] = np.inf
        S.add(f"fpe/subtract/{nm}/rows/inv", "fpe", "subtract", [Xi, Xi[:, :1]],
              lambda: flags_of(lambda: np.subtract(Xi, Xi[:, :1])))
        S.add(f"fpe/add/{nm}/rows/over", "fpe", "add", [Xo, Xo], lambda: flags_of(lambda: np.add(Xo, Xo)))
    # mixed and integer loops
    I = np.array([1, 0, -3, 0], dtype=np.int64)
    J = np.array([0, 0, 2, 5], dtype=np.int64)
    S.add("fpe/divide/int64", "fpe", "divide", [I, J], lambda: flags_of(lambda: np.divide(I, J)))
    S.add("fpe/divide/int64/nz", "fpe", "divide", [I, J + 1], lambda: flags_of(lambda: np.divide(I, J + 1)))
    F32 = np.array([3e38, 1.0], dtype=np.float32)
    F64 = np.array([10.0, 1e-320])
    S.add("fpe/multiply/f32xf64", "fpe", "multiply", [F32, F64], lambda: flags_of(lambda: np.multiply(F32, F64)))
    K = np.array([2 ** 62, 3], dtype=np.int64)
    S.add("fpe/add/int64", "fpe", "add", [K, K], lambda: flags_of(lambda: np.add(K, K)))
    # casts
    vals = [300.0, 3e9, -129.0, 1e20, np.nan, 255.5, -1.0, 2.0 ** 64, 2.0 ** 63, -2.0 ** 63, -9.3e18, 3.5e38, 1e-40,
            1e-46, np.inf, 3e38, 2.0 ** 31, -2.0 ** 31 - 1, -2.0 ** 31, 1.0e-39]
    for i, v in enumerate(vals):
        X = np.array([v, 1.0])
        for t in ("int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64", "bool", "float32"):
            S.add(f"fpe/cast/{i}/{t}", "fpe", "cast", [X], lambda: flags_of(lambda: X.astype(t)), {"dtype": t})
        if np.isfinite(v) and abs(v) < 3.4e38 or not np.isfinite(v):
            X32 = np.array([v, 1.0], dtype=np.float32)
            for t in ("int8", "int32", "uint32", "int64", "uint64"):
                S.add(f"fpe/cast32/{i}/{t}", "fpe", "cast", [X32], lambda: flags_of(lambda: X32.astype(t)), {"dtype": t})


def half_cases(S):
    """float16 (NumPy's HALF loops and npy_half conversions): arithmetic in
    float32 rounded back, maximum / minimum on the halves (first operand on
    ties), promotion with every dtype, casts both ways at the rounding and
    overflow / underflow boundaries, where, reductions per bufsize window
    (float32 accumulator stored back each window), accumulate, reduceat,
    argmax, IEEE flags.  Add / multiply inputs never pair two NaNs (which
    payload NumPy keeps then depends on the loop variant).  Own generator
    (seed 85)."""
    r10 = np.random.default_rng(85)
    h = np.float16

    def hvals(shape):
        x = (r10.standard_normal(shape) * r10.choice([1e-6, 1e-3, 1.0, 100.0, 3e4], shape)).astype(h)
        u = r10.random(shape)
        x[u < 0.03] = 0.0
        x[(u >= 0.03) & (u < 0.06)] = -0.0
        x[(u >= 0.06) & (u < 0.07)] = np.inf
        x[(u >= 0.07) & (u < 0.08)] = -np.inf
        x[(u >= 0.08) & (u < 0.09)] = np.nan
        v = x.view(np.uint16)
        m = (u >= 0.09) & (u < 0.1)
        v[m] = r10.integers(1, 1024, int(m.sum()))  # subnormals
        return x

    def partner(t, shape):
        if t == np.bool_:
            return r10.random(shape) < 0.5
        if np.dtype(t).kind in "iu":
            info = np.iinfo(t)
            return r10.integers(builtins_max(info.min, -70000), builtins_min(info.max, 70000), shape).astype(t)
        if t == h:
            return hvals(shape)
        return (r10.standard_normal(shape) * 100).astype(t)

    def flags_of(fn):
        got = []
        with np.errstate(all="call", call=lambda t, f: got.append(f)):
            fn()
        r = 0
        for f in got:
            r |= f
        return np.array([r], dtype=np.int64)

    ops = ("add", "subtract", "multiply", "true_divide", "maximum", "minimum", "less", "equal", "greater_equal")
    for t in (h, np.int8, np.uint8, np.bool_, np.int16, np.float32, np.float64, np.int64):
        nm = np.dtype(t).name
        a = hvals(400)
        b = partner(t, 400)
        for op in ops:
            if op == "subtract" and t == np.bool_:
                continue
            aa, bb = a, b
            if op in ("add", "multiply") and bb.dtype.kind == "f":
                bb = b.copy()
                bb[np.isnan(a.astype(np.float64)) & np.isnan(b.astype(np.float64))] = 1
            fn = getattr(np, op)
            S.add(f"half/{op}/{nm}", "binary", op, [aa, bb], lambda: fn(aa, bb))
            S.add(f"half/{op}/{nm}/rev", "binary", op, [bb, aa], lambda: fn(bb, aa))
    X = hvals((30, 40))
    col = hvals((30, 1))
    S.add("half/multiply/bcast", "binary", "multiply", [X, col], lambda: np.multiply(X, col))
    S.add("half/add/T", "binary", "add", [X.T, X[:1].T], lambda: np.add(X.T, X[:1].T))
    S.add("half/maximum/step", "binary", "maximum", [X[:, ::3], X[:, 1::3][:, :14]],
          lambda: np.maximum(X[:, ::3], X[:, 1::3][:, :14]))
    # casts
    x = hvals(600)
    for t in ("bool", "int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64", "float32", "float64"):
        S.add(f"half/cast/to/{t}", "cast", "cast", [x], lambda: x.astype(t), {"dtype": t})
    edge = np.array([1 + 2 ** -11 + 2 ** -30, 1 + 2 ** -11, 1 + 3 * 2 ** -11, 2 ** -25, 2 ** -25 * 1.0000001, 65519.99,
                     65520.0, 2 ** -24 * 1.5, 2 ** -24 * 2.5, 6.1e-5, 6.103515625e-05 * (1 - 2 ** -12), -0.0, 1e300,
                     -1e-300, np.inf, np.nan], dtype=np.float64)
    S.add("half/cast/from/float64/edge", "cast", "cast", [edge], lambda: edge.astype(h), {"dtype": "float16"})
    for t in (np.float32, np.float64):
        y = (r10.standard_normal(600) * r10.choice([1e-9, 1e-6, 1e-4, 1.0, 100.0, 6e4, 7e4], 600)).astype(t)
        y[::50] = np.nan
        if t == np.float32:
            yv = y.view(np.uint32)
            yv[1::97] = 0x7F800001 + r10.integers(0, 1 << 22, len(yv[1::97]), dtype=np.uint32)  # NaN payloads, signalling too
        S.add(f"half/cast/from/{np.dtype(t).name}", "cast", "cast", [y], lambda: y.astype(h), {"dtype": "float16"})
    for t in (np.bool_, np.int8, np.uint8, np.int16, np.int32, np.uint64):
        y = partner(t, 600)
        S.add(f"half/cast/from/{np.dtype(t).name}", "cast", "cast", [y], lambda: y.astype(h), {"dtype": "float16"})
    # where
    c = r10.random(300) < 0.5
    xh, yi, yf = hvals(300), partner(np.int8, 300), partner(np.float32, 300)
    S.add("half/where/f16,i8", "where", "where", [c, xh, yi], lambda: np.where(c, xh, yi))
    S.add("half/where/f16,f32", "where", "where", [c, xh, yf], lambda: np.where(c, xh, yf))
    # reductions
    for shape in ((100,), (20000,), (3, 9000), (300, 7)):
        g = (r10.standard_normal(shape) * 3).astype(h)
        gm = (1 + r10.standard_normal(shape) * 0.001).astype(h)
        for op in ("add", "multiply", "maximum", "minimum"):
            xx = gm if op == "multiply" else g
            for ax in [None] + list(range(len(shape))):
                S.add(f"half/reduce/{op}/{shape}/{ax}", "reduce", op, [xx], lambda: getattr(np, op).reduce(xx, axis=ax),
                      {"axis": ax})
            S.add(f"half/reduce/{op}/{shape}/init", "reduce", op, [xx],
                  lambda: getattr(np, op).reduce(xx, axis=0, initial=0.5), {"axis": 0, "initial": 0.5})
            S.add(f"half/accumulate/{op}/{shape}", "accumulate", op, [xx],
                  lambda: getattr(np, op).accumulate(xx, axis=len(shape) - 1), {"axis": len(shape) - 1})
            idx = sorted(set(r10.integers(0, shape[-1], 5).tolist()))
            S.add(f"half/reduceat/{op}/{shape}", "reduceat", op, [xx],
                  lambda: getattr(np, op).reduceat(xx, idx, axis=len(shape) - 1),
                  {"indices": [int(i) for i in idx], "axis": len(shape) - 1})
        S.add(f"half/argmax/{shape}", "argext", "argmax", [g], lambda: np.argmax(g, axis=len(shape) - 1),
              {"axis": len(shape) - 1})
    g = hvals(700)
    S.add("half/reduce/maximum/specials", "reduce", "maximum", [g], lambda: np.maximum.reduce(g), {"axis": 0})
    S.add("half/argmin/specials", "argext", "argmin", [g], lambda: np.argmin(g), {"axis": None})
    i8 = partner(np.int8, 5000)
    S.add("half/reduce/add/int8-dtype", "reduce", "add", [i8], lambda: np.add.reduce(i8, dtype=h),
          {"axis": 0, "dtype": "float16"})
    f32 = (r10.standard_normal(5000) * 3).astype(np.float32)
    S.add("half/reduce/add/f32-dtype", "reduce", "add", [f32], lambda: np.add.reduce(f32, dtype=h),
          {"axis": 0, "dtype": "float16"})
    g16 = (r10.standard_normal((40, 300)) * 3).astype(h)
    S.add("half/reduce/add/f16-dtype-f32", "reduce", "add", [g16], lambda: np.add.reduce(g16, axis=1, dtype=np.float32),
          {"axis": 1, "dtype": "float32"})
    S.add("half/cumsum/None", "accumulate", "cumsum", [g16], lambda: np.cumsum(g16), {"axis": None})
    # IEEE flags
    sp = np.array([0, -0.0, 1, -1, np.inf, -np.inf, np.nan, 65504, -65504, 6.1e-5, 2 ** -24, 2.0, 0.5, 3.0, 300, 1e-3],
                  dtype=h)
    for name in ("add", "subtract", "multiply", "divide"):
        fn = getattr(np, name)
        for i in range(len(sp)):
            a1 = sp[i:i + 1]
            S.add(f"half/fpe/{name}/row{i}", "fpe", name, [a1, sp], lambda: flags_of(lambda: fn(a1, sp)))
    for i, v in enumerate([70000.0, 65519.9, 65520.0, 1e-8, 2 ** -25, 2 ** -24 * 1.5, 3e-5, 1.0, -1e5]):
        for t in ("float32", "float64"):
            Y = np.array([v, 1.0], dtype=t)
            S.add(f"half/fpe/cast/{t}/{i}", "fpe", "cast", [Y], lambda: flags_of(lambda: Y.astype(h)),
                  {"dtype": "float16"})
    Yi = np.array([70000, 5], dtype=np.int32)
    S.add("half/fpe/cast/int32", "fpe", "cast", [Yi], lambda: flags_of(lambda: Yi.astype(h)), {"dtype": "float16"})
    for t in ("int8", "int32", "uint32", "int64", "uint64"):
        S.add(f"half/fpe/cast/to/{t}", "fpe", "cast", [sp], lambda: flags_of(lambda: sp.astype(t)), {"dtype": t})


builtins_max, builtins_min = max, min


def argext_cases(S, rng):
    for dt in (np.float64, np.float32, np.int64, np.int32, np.uint8, np.bool_):
        nm = np.dtype(dt).name
        for shp in ((1000,), (30, 50), (7, 300, 4)) + (((70000,),) if dt in (np.float64, np.int32) else ()):
            x = specials(dt, int(np.prod(shp)), rng).reshape(shp)
            if np.dtype(dt).kind == "f":
                x = np.nan_to_num(x, nan=0.0) if rng.random() < 0.5 else x
                x.reshape(-1)[rng.integers(0, x.size)] = x.max() if not np.isnan(x).any() else np.nan
            for vn, V in (("C", x), ("rev" if x.ndim == 1 else "T", x[::-1] if x.ndim == 1 else x.transpose())):
                for ax in [None] + list(range(V.ndim)):
                    for fn in ("argmax", "argmin"):
                        S.add(f"arg/{fn}/{nm}/{shp}/{vn}/{ax}", "argext", fn, [V],
                              lambda: getattr(np, fn)(V, axis=ax), {"axis": ax})
    x = np.array([1.0, 3.0, 3.0, np.nan, 5.0, np.nan])
    S.add("arg/argmax/nan_first", "argext", "argmax", [x], lambda: np.argmax(x), {"axis": None})
    x = np.array([-0.0, 0.0, 0.0])
    S.add("arg/argmax/zero_tie", "argext", "argmax", [x], lambda: np.argmax(x), {"axis": None})
    x = np.zeros((3, 4))
    S.add("arg/argmax/keepdims", "argext", "argmax", [x], lambda: np.argmax(x, axis=1, keepdims=True),
          {"axis": 1, "keepdims": True})
    S.add("error/argmax_empty", "argext", "argmax", [np.zeros((0,))], lambda: np.argmax(np.zeros((0,))), {"axis": None})


ALL_DTYPES = DTYPES + [np.float16]


def weak_scalar_cases(S):
    """NEP 50 weak Python scalars (ionp-py lib.rs:8238-8272
    extract_binary_pair -> scalar_against 8104): a Python bool / int /
    float against every dtype, on either side -- the scalar takes the
    array's dtype within its kind; an int outside it raises OverflowError
    (arithmetic) or compares by value (comparisons, also beyond 64 bits);
    true_divide's integer targets are float64; float scalars that overflow
    the array's float dtype.  Own generator (seed 91)."""
    r = np.random.default_rng(91)
    scal = [0, 1, -1, 7, 127, 128, 255, 256, -129, 40000, 2 ** 31, -2 ** 31 - 1, 2 ** 63 - 1, 2 ** 63, 2 ** 64 - 1,
            2 ** 64, -2 ** 63, -2 ** 63 - 1, 2 ** 70, -2 ** 70, 1.5, -0.0, 0.1, 1e300, 1e-50, float("nan"),
            float("inf"), True, False]
    ops = ["add", "subtract", "multiply", "true_divide", "maximum", "less", "equal", "greater_equal"]
    for dt in ALL_DTYPES:
        nm = np.dtype(dt).name
        x = specials(dt, 7, r) if dt != np.float16 else (r.standard_normal(7) * 100).astype(np.float16)
        for v in scal:
            for op in ops:
                if op == "subtract" and dt == np.bool_ and isinstance(v, bool):
                    continue
                f = getattr(np, op)
                S.add(f"weak/{op}/{nm}/{v!r}", "binary", op, [x, PyScalar(v)], lambda: f(x, v))
                if op in ("subtract", "less", "true_divide"):
                    S.add(f"weak/{op}/{v!r}/{nm}", "binary", op, [PyScalar(v), x], lambda: f(v, x))
    # both operands Python scalars
    for a, b in ((3, 4), (3, 4.5), (True, 2), (2 ** 63, -1), (1.5, True), (2 ** 64 - 1, 1)):
        for op in ("add", "less", "multiply"):
            f = getattr(np, op)
            S.add(f"weak/{op}/{a!r}/{b!r}", "binary", op, [PyScalar(a), PyScalar(b)], lambda: f(a, b))
    # where with a scalar branch
    c = r.random(9) < 0.5
    for dt in (np.int8, np.uint8, np.float32, np.float64, np.int64):
        y = specials(dt, 9, r)
        for v in (1, 300, -1, 2.5, True):
            S.add(f"weak/where/{np.dtype(dt).name}/{v!r}", "where", "where", [c, PyScalar(v), y],
                  lambda: np.where(c, v, y))


def dtype_arg_cases(S):
    """dtype= (and casting=) on Ufunc.__call__ (ionp-py lib.rs:9568-9804,
    resolve_binary_output_dtype ufunc.rs:1031): the inputs are cast into the
    loop of the requested output dtype first; rejected casts raise
    UFuncTypeError, a dtype no loop outputs raises TypeError; weak scalars
    are converted straight into the loop dtype.  Own generator (seed 92)."""
    r = np.random.default_rng(92)
    ops = ["add", "subtract", "multiply", "true_divide", "maximum", "less", "equal"]
    pairs = [(np.float64, np.float64), (np.float32, np.float64), (np.int32, np.int32), (np.int64, np.uint64),
             (np.int8, np.uint8), (np.bool_, np.bool_), (np.float16, np.int16), (np.int16, np.float32),
             (np.uint32, np.int8), (np.float64, np.int64)]
    for da, db in pairs:
        x = specials(da, 30, r).reshape(5, 6) if da != np.float16 else (r.standard_normal((5, 6)) * 50).astype(da)
        y = specials(db, 6, r) if db != np.float16 else (r.standard_normal(6) * 50).astype(db)
        for op in ops:
            f = getattr(np, op)
            for req in ALL_DTYPES:
                for casting in ("same_kind", "unsafe"):
                    if casting == "unsafe" and op not in ("add", "multiply"):
                        continue
                    if casting == "unsafe" and req == np.uint32 and "f" in (np.dtype(da).kind, np.dtype(db).kind):
                        continue  # float -> uint32 out of range: platform-defined in NumPy (unpinned)
                    rn = np.dtype(req).name
                    S.add(f"dtype/{op}/{np.dtype(da).name}-{np.dtype(db).name}/{rn}/{casting}", "binary", op, [x, y],
                          lambda: f(x, y, dtype=req, casting=casting), {"dtype": rn, "casting": casting})
    # weak scalars with dtype=
    x8 = np.array([1, -3, 100], np.int8)
    u8 = np.array([1, 3, 200], np.uint8)
    f32 = np.array([1.5, -2.25, 3e38], np.float32)
    for arr, v, req in ((x8, 300, np.int16), (u8, -1, np.int16), (u8, -1, np.uint16), (f32, 0.1, np.float64),
                        (x8, 1.5, np.int64), (x8, 2 ** 40, np.int16), (f32, 2, np.int64), (x8, 3, np.float32),
                        (f32, True, np.float32), (u8, 2 ** 70, np.int64)):
        rn = np.dtype(req).name
        for op in ("add", "multiply"):
            f = getattr(np, op)
            S.add(f"dtype/{op}/weak/{arr.dtype.name}/{v!r}/{rn}", "binary", op, [arr, PyScalar(v)],
                  lambda: f(arr, v, dtype=req), {"dtype": rn})


def where3_cases(S):
    """where with three-way broadcasting at rank 3 and 4: size-1 axes spread
    over condition, x and y, F-ordered and stepped / reversed operands, an
    integer condition (truthiness), mixed dtypes.  Own generator (seed 93)."""
    r = np.random.default_rng(93)
    for k in range(24):
        rank = 3 + k % 2
        shape = [int(v) for v in r.integers(2, 9, rank)]
        shapes = [list(shape) for _ in range(3)]
        for ax in range(rank):
            shapes[int(r.integers(0, 3))][ax] = 1
        dx, dy = [(np.float64, np.float32), (np.int64, np.uint8), (np.float32, np.int32), (np.int8, np.float64),
                  (np.bool_, np.int16), (np.float16, np.float32)][k % 6]
        ci = r.integers(0, 3, shapes[0]).astype(np.int32 if k % 5 == 0 else np.int8)
        c = ci if k % 5 == 0 else ci > 0
        x = specials(dx, int(np.prod(shapes[1])), r).reshape(shapes[1]) if dx != np.float16 else \
            (r.standard_normal(shapes[1]) * 10).astype(dx)
        y = specials(dy, int(np.prod(shapes[2])) * 2, r).reshape(shapes[2][:-1] + [shapes[2][-1] * 2])
        if k % 3 == 0:
            c = np.asfortranarray(c)
        if k % 4 == 1:
            x = np.asfortranarray(x)[::-1]
        ys = y[..., ::2] if k % 2 else y[..., 1::2][..., ::-1]
        S.add(f"where3/{k}/rank{rank}", "where", "where", [c, x, ys], lambda: np.where(c, x, ys))


def reduce_gap_cases(S):
    """float16 where= reductions (the HALF loop per maximal True run of a
    window: float32 accumulator stored back at the run's end;
    reduce_axis_masked_f16_widen, ufunc.rs:7414), maximum / minimum.reduce
    with a dtype= other than a float input's (NumPy's buffered cast: the
    loop reads the cast values; dtype rules ufunc.rs:2080-2163), float16
    into a float32 / float64 loop with a mask.  Own generator (seed 94)."""
    r = np.random.default_rng(94)
    h = np.float16
    for shp, ax in (((20000,), 0), ((300, 50), 1), ((50, 300), 0), ((3, 9000), 1), ((9000, 3), 0),
                    ((40, 30, 20), (0, 2)), ((17000,), None)):
        for v in range(2):
            x = (r.standard_normal(shp) * [1, 1e3][v]).astype(h)
            if v:
                x.reshape(-1)[[3, 50]] = [np.nan, -0.0]
            xm = (1 + 0.002 * r.standard_normal(shp)).astype(h)
            for p in (0.6, 0.05):
                m = r.random(shp) < p
                for op in ("add", "multiply", "maximum", "minimum"):
                    xi = xm if op == "multiply" else x
                    init = 1.5 if op in ("maximum", "minimum") else (-3.25 if v else None)
                    kw = {} if init is None else {"initial": init}
                    args = {"axis": list(ax) if isinstance(ax, tuple) else ax, "where": True, **kw}
                    S.add(f"gap/half_where/{op}/{shp}/{v}/{p}", "reduce", op, [xi, m],
                          lambda: getattr(np, op).reduce(xi, axis=ax, where=m, **kw), args)
                    if xi.ndim > 1:
                        S.add(f"gap/half_where/{op}/{shp}/{v}/{p}/T", "reduce", op, [xi.T, m.T],
                              lambda: getattr(np, op).reduce(xi.T, axis=ax, where=m.T, **kw), args)
    for shp, ax in (((20000,), 0), ((300, 50), 1), ((50, 300), 0), ((9000, 3), 0), ((40, 30, 20), (0, 2))):
        for src in (np.float64, np.float32, np.float16, np.int32, np.uint8):
            x = (r.standard_normal(shp) * 100).astype(src)
            if np.dtype(src).kind == "f":
                x.reshape(-1)[r.integers(0, x.size, 3)] = [np.nan, -0.0, 0.0]
            axv = list(ax) if isinstance(ax, tuple) else ax
            for dt in (np.float32, np.float64):
                if dt == src:
                    continue
                dn = np.dtype(dt).name
                for op in ("maximum", "minimum") + (("add",) if src == np.float16 else ()):
                    for vn, xv in (("C", x), ("T", x.T if x.ndim > 1 else x[::-2])):
                        S.add(f"gap/dtype/{op}/{np.dtype(src).name}->{dn}/{shp}/{vn}", "reduce", op, [xv],
                              lambda: getattr(np, op).reduce(xv, axis=ax, dtype=dt), {"axis": axv, "dtype": dn})
                        m = r.random(xv.shape) < 0.7
                        init = {"maximum": -1e9, "minimum": 1e9, "add": None}[op]
                        kw = {} if init is None else {"initial": init}
                        S.add(f"gap/dtype_where/{op}/{np.dtype(src).name}->{dn}/{shp}/{vn}", "reduce", op, [xv, m],
                              lambda: getattr(np, op).reduce(xv, axis=ax, dtype=dt, where=m, **kw),
                              {"axis": axv, "dtype": dn, "where": True, **kw})
    # float input into an integer / bool loop: the reduce family casts
    # unsafely by default (NumPy's C cast per element; out-of-range values
    # take the x86 conversion's result)
    for src in (np.float64, np.float32, np.float16):
        x = (r.standard_normal((40, 300)) * 100).astype(src)
        x[3, 5] = 1e10 if src != np.float16 else 6e4
        x[4, 7] = -1e10 if src != np.float16 else -6e4
        x[5, 9] = -0.7
        sn = np.dtype(src).name
        for dt in (np.int64, np.int32, np.int8, np.uint8, np.uint64, np.bool_):
            dn = np.dtype(dt).name
            for ax in (0, 1, None):
                for op in ("add", "maximum", "multiply"):
                    S.add(f"gap/f2i/{op}/{sn}->{dn}/{ax}", "reduce", op, [x],
                          lambda: getattr(np, op).reduce(x, axis=ax, dtype=dt), {"axis": ax, "dtype": dn})
                m = r.random(x.shape) < 0.5
                S.add(f"gap/f2i_where/add/{sn}->{dn}/{ax}", "reduce", "add", [x, m],
                      lambda: np.add.reduce(x, axis=ax, dtype=dt, where=m), {"axis": ax, "dtype": dn, "where": True})
            for ax in (0, 1):
                S.add(f"gap/f2i_acc/add/{sn}->{dn}/{ax}", "accumulate", "add", [x],
                      lambda: np.add.accumulate(x, axis=ax, dtype=dt), {"axis": ax, "dtype": dn})
                S.add(f"gap/f2i_rat/add/{sn}->{dn}/{ax}", "reduceat", "add", [x],
                      lambda: np.add.reduceat(x, [0, 5, 2, 20], axis=ax, dtype=dt),
                      {"axis": ax, "indices": [0, 5, 2, 20], "dtype": dn})


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def big_workloads():
    """Digests of the full-size workloads (inputs regenerated from seeds)."""
    out = {}
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, (1024, 1024))
    b = rng.uniform(-1, 1, 1024)
    C = np.add(A, b)
    out["cfg1"] = {"C": sha(C), "sum0": sha(np.add.reduce(C, 0)), "sum1": sha(np.add.reduce(C, 1)),
                   "sumNone": float(np.add.reduce(C, axis=None)).hex()}
    X = np.random.default_rng(1).standard_normal((8192, 8192))
    c = np.random.default_rng(2).uniform(0.5, 2, (8192, 1))
    X32 = X.astype(np.float32)
    c32 = c.astype(np.float32)
    out["cfg2"] = {"mul_f64": sha(X * c), "mul_f32": sha(X32 * c32),
                   "view_f32": sha(np.multiply(X32.T[::2], c32[::2])),
                   "view_f64": sha(np.multiply(X.T[::2], c[::2]))}
    del X, X32
    I = np.random.default_rng(3).integers(-2 ** 62, 2 ** 62, (8192, 8192))
    j = np.random.default_rng(8).integers(-2 ** 62, 2 ** 62, (8192, 1))
    out["cfg2"]["add_i64"] = sha(I + j)
    del I
    rng4 = np.random.default_rng(4)
    x = rng4.uniform(-1, 1, 2 ** 28)
    out["cfg3"] = {"sum_f64": float(np.add.reduce(x)).hex()}
    xn = x.copy()
    xn[np.random.default_rng(5).random(2 ** 28) < 1e-6] = np.nan
    out["cfg3"]["max_nan"] = hex(np.maximum.reduce(xn).view(np.uint64))
    out["cfg3"]["argmax_nan"] = int(np.argmax(xn))
    out["cfg3"]["max"] = hex(np.maximum.reduce(x).view(np.uint64))
    out["cfg3"]["argmax"] = int(np.argmax(x))
    del x, xn
    F = np.random.default_rng(6).uniform(-1, 1, (4096, 65536)).astype(np.float32)
    out["cfg3"]["f32_sum0"] = sha(np.add.reduce(F, axis=0))
    out["cfg3"]["f32_sum1"] = sha(np.add.reduce(F, axis=1))
    del F
    I32 = np.random.default_rng(7).integers(-1000, 1000, 2 ** 28, dtype=np.int32)
    out["cfg3"]["i32_sum"] = int(np.sum(I32))
    return out


def main():
    rng = np.random.default_rng(20260731)  # the reference corpus seed (corpus.py:22)
    S = Store()
    binary_cases(S, rng)
    where_cast_cases(S, rng)
    reduce_cases(S, rng)
    argext_cases(S, rng)
    accumulate_cases(S)
    prod_cases(S)
    where_reduce_cases(S)
    reduceat_cases(S)
    fpe_cases(S)
    half_cases(S)
    weak_scalar_cases(S)
    dtype_arg_cases(S)
    where3_cases(S)
    reduce_gap_cases(S)
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **S.arrays)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(S.cases, f)
    print(f"{len(S.cases)} cases, {sum(a.nbytes for a in S.arrays.values()) / 1e6:.1f} MB raw")
    if "--no-big" not in sys.argv:
        with open(os.path.join(HERE, "big.json"), "w") as f:
            json.dump(big_workloads(), f, indent=1)


if __name__ == "__main__":
    main()
