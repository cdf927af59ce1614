"""Full-size workloads (inputs regenerated from their seeds) compared
bit-for-bit with NumPy 2.2.6 digests recorded in tests/golden/big.json, and
reductions at the folder / chain / window boundaries compared with the C
oracle (itself pinned to NumPy 2.2.6 by the golden vectors)."""
import hashlib

import numpy as np
import pytest
import torch

import oracle as O
from fixture_cases import big

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def digests():
    d = big()
    if d is None:
        pytest.fail("tests/golden/big.json missing")
    return d


def test_cfg1_broadcast_add_and_reduce(ap, digests):
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, (1024, 1024))
    b = rng.uniform(-1, 1, 1024)
    C = ap.add(ap.asarray(A), ap.asarray(b))
    d = digests["cfg1"]
    assert sha(C.get()) == d["C"]
    assert sha(ap.add.reduce(C, axis=0).get()) == d["sum0"]
    assert sha(ap.add.reduce(C, axis=1).get()) == d["sum1"]
    assert float(ap.add.reduce(C, axis=None).get()).hex() == d["sumNone"]


def test_cfg2_broadcast_multiply(ap, digests):
    d = digests["cfg2"]
    X = np.random.default_rng(1).standard_normal((8192, 8192))
    c = np.random.default_rng(2).uniform(0.5, 2, (8192, 1))
    Xd, cd = ap.asarray(X), ap.asarray(c)
    assert sha(ap.multiply(Xd, cd).get()) == d["mul_f64"]
    assert sha(ap.multiply(Xd.T[::2], cd[::2]).get()) == d["view_f64"]
    del Xd
    X32d, c32d = ap.asarray(X.astype(np.float32)), ap.asarray(c.astype(np.float32))
    del X
    assert sha(ap.multiply(X32d, c32d).get()) == d["mul_f32"]
    assert sha(ap.multiply(X32d.T[::2], c32d[::2]).get()) == d["view_f32"]
    del X32d
    I = np.random.default_rng(3).integers(-2 ** 62, 2 ** 62, (8192, 8192))
    j = np.random.default_rng(8).integers(-2 ** 62, 2 ** 62, (8192, 1))
    assert sha(ap.add(ap.asarray(I), ap.asarray(j)).get()) == d["add_i64"]
    torch.cuda.empty_cache()


def test_cfg3_reductions(ap, digests):
    d = digests["cfg3"]
    x = np.random.default_rng(4).uniform(-1, 1, 2 ** 28)
    xd = ap.asarray(x)
    assert float(ap.add.reduce(xd).get()).hex() == d["sum_f64"]
    assert hex(ap.maximum.reduce(xd).get().view(np.uint64)) == d["max"]
    assert int(ap.argmax(xd).get()) == d["argmax"]
    x[np.random.default_rng(5).random(2 ** 28) < 1e-6] = np.nan
    xd = ap.asarray(x)
    del x
    assert hex(ap.maximum.reduce(xd).get().view(np.uint64)) == d["max_nan"]
    assert int(ap.argmax(xd).get()) == d["argmax_nan"]
    del xd
    F = ap.asarray(np.random.default_rng(6).uniform(-1, 1, (4096, 65536)).astype(np.float32))
    assert sha(ap.add.reduce(F, axis=0).get()) == d["f32_sum0"]
    assert sha(ap.add.reduce(F, axis=1).get()) == d["f32_sum1"]
    del F
    I = ap.asarray(np.random.default_rng(7).integers(-1000, 1000, 2 ** 28, dtype=np.int32))
    s = ap.sum(I)
    assert s.dtype == np.int64 and int(s.get()) == d["i32_sum"]
    torch.cuda.empty_cache()


def test_single_launch_folds_repeat(ap):
    """Reductions that fold per-workgroup partials (maximum/minimum over
    contiguous windows, argmax/argmin over 65536-element segments, float
    row sums without folder workgroups) reuse the library's scratch memory;
    interleaved repeats must each match NumPy."""
    rng = np.random.default_rng(12)
    x = rng.standard_normal(3 * 2 ** 20 + 5)
    R = rng.standard_normal((300, 2 * 8192 + 3)).astype(np.float32)
    xd, Rd = ap.asarray(x), ap.asarray(R)
    exp = {"max": O.reduce("maximum", x, None), "min": O.reduce("minimum", x, None),
           "argmax": int(O.argmax(x)), "argmin": int(O.argmin(x)),
           "rows": np.asarray(O.reduce("add", R, 1)), "rowmax": np.asarray(O.reduce("maximum", R, 1))}
    for _ in range(6):
        assert ap.maximum.reduce(xd).get() == exp["max"]
        assert int(ap.argmax(xd).get()) == exp["argmax"]
        assert np.array_equal(ap.add.reduce(Rd, axis=1).get().view(np.uint32), exp["rows"].view(np.uint32))
        assert ap.minimum.reduce(xd).get() == exp["min"]
        assert int(ap.argmin(xd).get()) == exp["argmin"]
        assert np.array_equal(ap.maximum.reduce(Rd, axis=1).get(), exp["rowmax"])


def test_folder_sum_special_values_and_repeat(ap):
    """One-output float sums long enough for the in-launch folder (more than
    8 windows, with a partial last window): NaN payloads, infinities of both
    signs and overflow must give NumPy's bits, and every call must leave the
    partial slots ready for the next one (repeats, alternating float32 and
    float64)."""
    rng = np.random.default_rng(13)
    base = rng.standard_normal(2 ** 20 + 77)
    nan1 = np.array([0x7FF8000000000123], dtype=np.uint64).view(np.float64)[0]
    cases = []
    x = base.copy()
    cases.append(x)
    x = base.copy(); x[5000] = nan1; cases.append(x)
    x = base.copy(); x[900001] = nan1; x[1000000] = np.inf; cases.append(x)
    x = base.copy(); x[10] = np.inf; x[700000] = -np.inf; cases.append(x)
    x = base.copy(); x[123456] = np.inf; cases.append(x)
    x = base.copy(); x[:3] = 1.7e308; cases.append(x)
    for _ in range(2):
        for h in cases:
            for dt in (np.float64, np.float32):
                hd = h.astype(dt)
                exp = np.atleast_1d(O.reduce("add", hd, None)).view(np.uint64 if dt == np.float64 else np.uint32)
                got = np.atleast_1d(ap.add.reduce(ap.asarray(hd)).get()).view(exp.dtype)
                assert got[0] == exp[0], (dt, hex(int(got[0])), hex(int(exp[0])))


def test_narrow_wave_chains(ap):
    """Few outputs with long sequential chains (a wavefront per output, the
    chain read back from LDS): block tails, a NaN payload, infinities, both
    float widths, against NumPy's bits."""
    rng = np.random.default_rng(14)
    nan1 = np.array([0x7FF8000000000321], dtype=np.uint64).view(np.float64)[0]
    for shape in ((300, 5), (256, 3), (257, 2), (70001, 3), (4099, 7)):
        for dt in (np.float64, np.float32):
            h = rng.standard_normal(shape).astype(dt)
            if shape[0] > 1000:
                h[1000, 0] = nan1
                h[shape[0] // 2, 1] = np.inf
                h[shape[0] - 1, 1] = -np.inf
            exp = np.asarray(O.reduce("add", h, 0))
            got = ap.add.reduce(ap.asarray(h), axis=0).get()
            assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (shape, dt)


def test_product_chains(ap):
    """multiply.reduce chains long enough for a wavefront per output (the
    factors read back from LDS): 1-D and axis-0 layouts, block tails, a NaN,
    a where= mask, against NumPy's bits."""
    rng = np.random.default_rng(15)
    for shape, axis in (((70001,), 0), ((5000, 3), 0), ((3, 4099), 1), ((257, 2), 0)):
        for dt in (np.float64, np.float32):
            h = (1 + 1e-3 * rng.standard_normal(shape)).astype(dt)
            if shape[0] == 5000:
                h[2500, 1] = np.nan
            exp = np.asarray(O.reduce("multiply", h, axis))
            got = ap.multiply.reduce(ap.asarray(h), axis=axis).get()
            assert np.array_equal(np.atleast_1d(got).view(np.uint8), np.atleast_1d(exp).view(np.uint8)), (shape, dt)
            m = rng.random(shape) < 0.7
            exp = np.asarray(O.reduce("multiply", h, axis, where=m))
            got = ap.multiply.reduce(ap.asarray(h), axis=axis, where=ap.asarray(m)).get()
            assert np.array_equal(np.atleast_1d(got).view(np.uint8), np.atleast_1d(exp).view(np.uint8)), (shape, dt, "where")


def test_streaming_sum_misaligned_views(ap):
    """The streaming sum reads 16-byte vectors: views whose rows start off a
    16-byte boundary (an element offset, odd row lengths) must still give
    NumPy's bits, with and without folder workgroups."""
    rng = np.random.default_rng(16)
    for dt in (np.float64, np.float32):
        base = rng.standard_normal(2 ** 20 + 7).astype(dt)
        d = ap.asarray(base)
        for k in (1, 2, 3):
            exp = O.reduce("add", base[k:], None)
            got = ap.add.reduce(d[k:]).get()
            assert np.atleast_1d(got).view(np.uint8).tobytes() == np.atleast_1d(exp).view(np.uint8).tobytes(), (dt, k)
        R = rng.standard_normal((70, 3 * 8192 + 5)).astype(dt)
        Rd = ap.asarray(R)
        for sl in (np.s_[:, :], np.s_[:, 1:], np.s_[1:, 3:]):
            exp = np.asarray(O.reduce("add", R[sl], 1))
            got = ap.add.reduce(Rd[sl], axis=1).get()
            assert np.array_equal(got.view(np.uint8), exp.view(np.uint8)), (dt, sl)


def test_transposed_operand_tiles(ap):
    """One operand contiguous along another axis than the output's (A + B.T):
    full 128 x 128 / 32 x 128 vector tiles plus partial edge tiles, both
    operand orders, comparisons (bool output), a broadcast row, a 3-D batch
    and a misaligned view (the scalar tile kernel), against the oracle."""
    rng = np.random.default_rng(17)
    for dt in (np.float32, np.float64, np.int32, np.int64):
        kind = np.dtype(dt).kind
        mk = (lambda s: rng.standard_normal(s).astype(dt)) if kind == "f" else \
            (lambda s: rng.integers(-10 ** 6, 10 ** 6, s).astype(dt))
        A, B = mk((300, 1000)), mk((1000, 300))
        row = mk((1, 1000))
        C, D = mk((3, 260, 129)), mk((3, 129, 260))
        Ad, Bd, rowd, Cd, Dd = (ap.asarray(v) for v in (A, B, row, C, D))
        cases = [("add", A, B.T, Ad, Bd.T), ("subtract", B.T, A, Bd.T, Ad), ("less", A, B.T, Ad, Bd.T),
                 ("maximum", row, B.T, rowd, Bd.T), ("multiply", C, D.transpose(0, 2, 1), Cd, Dd.transpose(0, 2, 1)),
                 ("add", A[:, 1:], B.T[:, 1:], Ad[:, 1:], Bd.T[:, 1:])]
        if kind == "f":
            cases.append(("true_divide", A, B.T, Ad, Bd.T))
        for op, a, b, ad, bd in cases:
            want = O.binary(op, a, b)
            got = getattr(ap, op)(ad, bd).get()
            assert got.strides == want.strides and got.tobytes() == want.tobytes(), (dt, op, a.shape)


def test_transposed_operand_any_call(ap):
    """Calls with an input contiguous along another axis than the output's
    that the typed tile kernels do not take -- where, casts into a C-ordered
    output, mixed-dtype arithmetic and comparisons (int64 against uint64
    exactly) -- staged through LDS in 64 x 64 tiles: partial edge tiles, a
    3-D batch, a reversed view and broadcast operands, against the oracle;
    and the IEEE-flag variant (overflow raised, a clean call silent)."""
    rng = np.random.default_rng(23)
    f8 = rng.standard_normal((300, 1000))
    f4 = rng.standard_normal((1000, 300)).astype(np.float32)
    i8 = rng.integers(-10 ** 9, 10 ** 9, (300, 1000))
    u8 = rng.integers(0, 2 ** 63, (1000, 300), dtype=np.uint64)
    i4 = rng.integers(-10 ** 6, 10 ** 6, (1000, 300)).astype(np.int32)
    cond = rng.random((300, 1000)) < 0.5
    condT = rng.random((1000, 300)) < 0.5
    f8b = rng.standard_normal((1000, 300))
    c3 = rng.random((3, 130, 70)) < 0.5
    x3 = rng.standard_normal((3, 70, 130))
    y3 = rng.standard_normal((3, 130, 70)).astype(np.float32)
    D = {k: ap.asarray(v) for k, v in dict(f8=f8, f4=f4, i8=i8, u8=u8, i4=i4, cond=cond, condT=condT, f8b=f8b,
                                            c3=c3, x3=x3, y3=y3).items()}
    where_cases = [
        (cond, f8, f8b.T, D["cond"], D["f8"], D["f8b"].T),
        (condT.T, f8, f8b.T, D["condT"].T, D["f8"], D["f8b"].T),
        (cond, f4.T, np.float32(2.5), D["cond"], D["f4"].T, 2.5),  # a weak Python float: float32
        (cond, i8, i4.T, D["cond"], D["i8"], D["i4"].T),
        (c3, x3.transpose(0, 2, 1), y3, D["c3"], D["x3"].transpose(0, 2, 1), D["y3"]),
        (cond[::-1], f8[::-1], f8b.T[::-1], D["cond"][::-1], D["f8"][::-1], D["f8b"].T[::-1]),
    ]
    for c, x, y, cd, xd, yd in where_cases:
        want = O.where(c, x, y)
        got = ap.where(cd, xd, yd).get()
        assert got.dtype == want.dtype and got.strides == want.strides and got.tobytes() == want.tobytes(), \
            (np.shape(c), np.asarray(x).dtype)
    for src, srcd in ((f4, D["f4"]), (f8b, D["f8b"]), (u8, D["u8"]), (condT, D["condT"])):
        for dt in (np.float64, np.float32, np.int64, np.int8, np.uint16, np.bool_):
            want = np.ascontiguousarray(O.cast(src.T, dt))
            got = srcd.T.astype(dt, order="C").get()
            assert got.dtype == want.dtype and got.tobytes() == want.tobytes(), (src.dtype, dt)
    bin_cases = [("add", f8, f4.T, D["f8"], D["f4"].T), ("multiply", i8, i4.T, D["i8"], D["i4"].T),
                 ("less", i8, u8.T, D["i8"], D["u8"].T), ("greater_equal", f4.T, i8, D["f4"].T, D["i8"]),
                 ("subtract", x3.transpose(0, 2, 1)[:, :, 1:], y3[:, :, 1:], D["x3"].transpose(0, 2, 1)[:, :, 1:],
                  D["y3"][:, :, 1:])]
    for op, a, b, ad, bd in bin_cases:
        want = O.binary(op, a, b)
        got = getattr(ap, op)(ad, bd).get()
        assert got.strides == want.strides and got.tobytes() == want.tobytes(), (op, a.dtype, b.dtype)
    big = np.full((200, 130), 3e38, dtype=np.float32)
    huge = np.full((130, 200), 1e300)
    assert O.binary_fpe("multiply", huge, big.T) & 2 and O.binary_fpe("add", f8, f4.T) == 0
    bigd, huged = ap.asarray(big), ap.asarray(huge)
    with ap.errstate(over="raise", invalid="ignore", under="ignore", divide="ignore"):
        with pytest.raises(FloatingPointError, match="overflow"):
            ap.multiply(huged, bigd.T)
        # the typed cast into another order with flags tracked
        with pytest.raises(FloatingPointError, match="overflow"):
            huged.T.astype(np.float32, order="C")
        got_cast = D["f8b"].T.astype(np.float32, order="C").get()
    assert O.cast_fpe(huge.T, np.float32) & 2 and O.cast_fpe(f8b.T, np.float32) == 0
    assert got_cast.tobytes() == np.ascontiguousarray(O.cast(f8b.T, np.float32)).tobytes()
    with ap.errstate(over="raise", invalid="ignore", under="ignore", divide="ignore"):
        got = ap.add(D["f8"], D["f4"].T).get()
    assert got.tobytes() == O.binary("add", f8, f4.T).tobytes()


def test_transposed_operand_any_call_edges(ap):
    """The LDS tile kernel's other variants and limits, against the oracle:
    where bit selects of 2- and 4-byte values, a transposed condition in a
    3-D batch, typed casts from bool and from float64 into uint8 (cvt edge
    values), int32 / int64 maximum and minimum, a staged operand reversed
    along the output's contiguous dim, and the output-contiguous extent at
    the kernel's 32-element limit and just below it (the generic kernel)."""
    rng = np.random.default_rng(29)
    for dt in (np.int16, np.float32, np.int32):
        c = rng.random((130, 257)) < 0.5
        x = rng.integers(-1000, 1000, (130, 257)).astype(dt)
        y = rng.integers(-1000, 1000, (257, 130)).astype(dt)
        got = ap.where(ap.asarray(c), ap.asarray(x), ap.asarray(y).T).get()
        want = O.where(c, x, y.T)
        assert got.dtype == want.dtype and got.strides == want.strides and got.tobytes() == want.tobytes(), dt
    cT = rng.random((3, 90, 70)) < 0.5
    x3 = rng.standard_normal((3, 70, 90))
    y3 = rng.standard_normal((3, 70, 90))
    got = ap.where(ap.asarray(cT).transpose(0, 2, 1), ap.asarray(x3), ap.asarray(y3)).get()
    want = O.where(cT.transpose(0, 2, 1), x3, y3)
    assert got.strides == want.strides and got.tobytes() == want.tobytes()
    b = rng.random((300, 200)) < 0.5
    v = np.concatenate([rng.uniform(-300, 300, 59990), [np.nan, np.inf, -np.inf, 255.5, -0.5, 256.0, 1e20, -1e20,
                                                         0.0, -0.0]]).reshape(300, 200)
    for src, dt in ((b, np.float64), (b, np.int32), (v, np.uint8), (v, np.int16)):
        want = np.ascontiguousarray(O.cast(src.T, dt))
        with ap.errstate(all="ignore"):
            got = ap.asarray(src).T.astype(dt, order="C").get()
        assert got.dtype == want.dtype and got.tobytes() == want.tobytes(), (src.dtype, dt)
    i4 = rng.integers(-2 ** 31, 2 ** 31, (200, 300)).astype(np.int32)
    i8 = rng.integers(-2 ** 40, 2 ** 40, (300, 200))
    for op in ("maximum", "minimum", "subtract"):
        want = O.binary(op, i4, i8.T)
        got = getattr(ap, op)(ap.asarray(i4), ap.asarray(i8).T).get()
        assert got.strides == want.strides and got.tobytes() == want.tobytes(), op
    f8 = rng.standard_normal((300, 200))
    f4 = rng.standard_normal((200, 300)).astype(np.float32)
    for a, bt, A, BT in ((f8, f4.T[:, ::-1], ap.asarray(f8), ap.asarray(f4).T[:, ::-1]),
                         (f8[::-1], f4.T[::-1], ap.asarray(f8)[::-1], ap.asarray(f4).T[::-1])):
        want = O.binary("add", a, bt)
        got = ap.add(A, BT).get()
        assert got.strides == want.strides and got.tobytes() == want.tobytes()
    for n0 in (32, 31):
        a = rng.standard_normal((500, n0))
        bb = rng.standard_normal((n0, 500)).astype(np.float32)
        want = O.binary("multiply", a, bb.T)
        got = ap.multiply(ap.asarray(a), ap.asarray(bb).T).get()
        assert got.strides == want.strides and got.tobytes() == want.tobytes(), n0
    # outputs whose rows are 16-byte multiples: full tiles stored as 16-byte
    # vectors (where of every value size, casts with the cvt edges and the
    # flag variant, mixed-dtype arithmetic; comparisons keep item stores)
    for dt in (np.int8, np.int16, np.float32, np.float64):
        c = rng.random((130, 256)) < 0.5
        x = rng.integers(-100, 100, (130, 256)).astype(dt)
        y = rng.integers(-100, 100, (256, 130)).astype(dt)
        got = ap.where(ap.asarray(c), ap.asarray(x), ap.asarray(y).T).get()
        want = O.where(c, x, y.T)
        assert got.strides == want.strides and got.tobytes() == want.tobytes(), dt
    va = np.concatenate([rng.uniform(-300, 300, 256 * 320 - 10), [np.nan, np.inf, -np.inf, 255.5, -0.5, 256.0, 1e20,
                                                                  -1e20, 0.0, -0.0]]).reshape(256, 320)
    ba = rng.random((256, 320)) < 0.5
    for src, dt in ((va, np.uint8), (va, np.bool_), (va, np.int16), (va, np.float32), (va, np.int64),
                    (ba, np.float64), (ba, np.uint8)):
        want = np.ascontiguousarray(O.cast(src.T, dt))
        with ap.errstate(all="ignore"):
            got = ap.asarray(src).T.astype(dt, order="C").get()
        assert got.dtype == want.dtype and got.tobytes() == want.tobytes(), (src.dtype, dt)
    hugea = np.full((256, 192), 1e300)
    assert O.cast_fpe(hugea.T, np.float32) & 2
    with ap.errstate(over="raise", invalid="ignore", under="ignore", divide="ignore"):
        with pytest.raises(FloatingPointError, match="overflow"):
            ap.asarray(hugea).T.astype(np.float32, order="C")
        got = ap.asarray(va[:, :192] / 1e3).T.astype(np.float32, order="C").get()
    assert got.tobytes() == np.ascontiguousarray(O.cast((va[:, :192] / 1e3).T, np.float32)).tobytes()
    # same-dtype float arithmetic with IEEE flags tracked (the typed tile
    # kernel with flags; untracked, ew_tiled_v takes these): partial tiles,
    # a clean call under raise, an overflow raised
    for dt in (np.float64, np.float32):
        p = rng.standard_normal((300, 1000)).astype(dt)
        qT = rng.standard_normal((1000, 300)).astype(dt)
        with ap.errstate(all="raise"):
            got = ap.subtract(ap.asarray(p), ap.asarray(qT).T).get()
            got_m = ap.multiply(ap.asarray(p), ap.asarray(qT).T).get()
        assert got.tobytes() == O.binary("subtract", p, qT.T).tobytes(), dt
        assert got_m.tobytes() == O.binary("multiply", p, qT.T).tobytes(), dt
        hp = np.full((256, 192), 1e30 if dt == np.float32 else 1e300, dtype=dt)
        assert O.binary_fpe("multiply", hp, hp.T.copy().T) & 2
        with ap.errstate(over="raise", invalid="ignore", under="ignore", divide="ignore"):
            with pytest.raises(FloatingPointError, match="overflow"):
                ap.multiply(ap.asarray(hp), ap.asarray(np.ascontiguousarray(hp.T)).T)
    ia = rng.integers(-2 ** 31, 2 ** 31, (256, 320)).astype(np.int32)
    la = rng.integers(-2 ** 40, 2 ** 40, (320, 256))
    fa = rng.standard_normal((320, 256)).astype(np.float32)
    for op, a, b, ad, bd in (("maximum", ia, la.T, ap.asarray(ia), ap.asarray(la).T),
                             ("add", ia, fa.T, ap.asarray(ia), ap.asarray(fa).T),
                             ("less", ia, la.T, ap.asarray(ia), ap.asarray(la).T)):
        want = O.binary(op, a, b)
        got = getattr(ap, op)(ad, bd).get()
        assert got.strides == want.strides and got.tobytes() == want.tobytes(), op


def test_blocked_accumulate(ap):
    """Scans over many fibers (the blocked kernel): along the contiguous axis
    and along the outer axis of C-ordered arrays, extents that are not
    multiples of the block shape, a NaN (the fiber redone with x86 NaN
    rules), integer wrap, against the oracle."""
    rng = np.random.default_rng(18)
    for shape in ((3000, 2100), (2100, 3000), (4, 700, 530)):
        for dt in (np.float64, np.float32, np.int64, np.int32):
            if np.dtype(dt).kind == "f":
                h = rng.standard_normal(shape).astype(dt)
                h.reshape(-1)[12345] = np.nan
            else:
                h = rng.integers(-2 ** 30, 2 ** 30, shape).astype(dt)
            d = ap.asarray(h)
            for ax in range(h.ndim):
                for op in ("add", "maximum") + (("multiply",) if np.dtype(dt).kind == "i" else ()):
                    want = O.accumulate(op, h, ax)
                    got = getattr(ap, op).accumulate(d, axis=ax).get()
                    assert got.tobytes() == want.tobytes(), (shape, dt, ax, op)
                    if op == "add":  # 8 and 32 fibers per workgroup (ionp_config.split)
                        for fpb in (8, 32):
                            o = getattr(ap, op).accumulate(d, axis=ax)
                            o._buf.fill_(0xAB)
                            rc = _lib_accumulate(ap, _lib_code(op), d, ax, o, split=fpb)
                            assert rc == 0 and o.get().tobytes() == want.tobytes(), (shape, dt, ax, op, fpb)


def _lib_code(op):
    from anionpy_amd import _lib
    return {"add": _lib.ADD, "multiply": _lib.MULTIPLY, "maximum": _lib.MAXIMUM}[op]


def _lib_accumulate(ap, code, x, axis, out, split=0):
    """ionp_accumulate through the C ABI with an explicit configuration."""
    import ctypes
    from anionpy_amd import _lib
    dx, do = x._desc(), out._desc()
    cfg = _lib.Config(0, 0, 0, split)
    rc = _lib.load().ionp_accumulate(code, ctypes.byref(dx), axis, ctypes.byref(do), ctypes.byref(cfg),
                                     _lib.stream_handle())
    import torch
    torch.cuda.synchronize()
    return rc


def _values(rng, shape, dt):
    """Random values of dtype dt with the edge cases of casts mixed in."""
    dt = np.dtype(dt)
    n = int(np.prod(shape))
    if dt.kind == "f":
        v = rng.standard_normal(n) * 1e3
        sp = np.array([np.nan, -np.nan, np.inf, -np.inf, 1e20, -1e20, -0.0, 0.5, -0.5, 2.0 ** 63, 2.0 ** 64, -1.0,
                       2.0 ** 31, -2.0 ** 31 - 1, 255.5, 65535.9, 3e38, -3e38, 1e-310])
        idx = rng.choice(n, size=min(n, 4 * len(sp)), replace=False)
        v[idx] = np.resize(sp, len(idx))
        with np.errstate(all="ignore"):
            return v.astype(dt).reshape(shape)
    if dt.kind == "b":
        return (rng.random(n) < 0.5).reshape(shape)
    info = np.iinfo(dt)
    return rng.integers(info.min, info.max, n, dtype=dt, endpoint=True).reshape(shape)


def test_rows_where_and_cast(ap):
    """where and casts over rows layouts (the vectorised where / cast
    kernels): every value dtype, every cast pair, NaN payloads and
    out-of-range values, broadcast rows / columns, ragged tails, a 3-D
    layout and a misaligned view (the generic kernel), byte-for-byte
    against the oracle; one large case of each streams through the
    nontemporal path."""
    rng = np.random.default_rng(23)
    dts = (np.bool_, np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64,
           np.float32, np.float64)
    for dt in dts:
        # contiguous (one coalesced dim with a ragged tail), aligned rows of
        # 1024 with broadcast columns / rows and a ragged 1000-wide view,
        # a 3-D layout; then layouts the vector path declines (a 1003
        # pitch, an odd element offset)
        cond, x, y = rng.random((300, 1003)) < 0.5, _values(rng, (300, 1003), dt), _values(rng, (300, 1003), dt)
        cond2, x2, y2 = rng.random((64, 1024)) < 0.5, _values(rng, (64, 1024), dt), _values(rng, (64, 1024), dt)
        col, row = _values(rng, (64, 1), dt), _values(rng, (1, 1024), dt)
        c3, x3, y3 = rng.random((3, 50, 776)) < 0.3, _values(rng, (3, 50, 776), dt), _values(rng, (1, 50, 1), dt)
        cases = [(cond, x, y), (cond2, col, row), (cond2[:1], x2, col), (cond2, x2, y2[:1]),
                 (cond2[:, :1000], x2[:, :1000], row[:, :1000]), (c3, x3, y3),
                 (cond2, x2[:, ::-1], y2), (cond2[:, ::-1], col, y2[:, ::-1]), (cond2[:, :100], x2[:, 99::-1], y2[:, :100]),
                 (cond[:, :1000], x[:, :1000], y[:, :1000]), (cond[:, 1:], x[:, 1:], y[:, 1:])]
        for c, a, b in cases:
            want = O.where(c, a, b)
            got = ap.where(ap.asarray(c), ap.asarray(a), ap.asarray(b)).get()
            assert got.dtype == want.dtype and got.strides == want.strides, (dt, a.shape)
            assert got.tobytes() == want.tobytes(), ("where", dt, c.shape, a.shape, b.shape)
    for src in dts:
        a, b = _values(rng, (200, 1001), src), _values(rng, (64, 1024), src)
        ad, bd = ap.asarray(a), ap.asarray(b)
        for to in dts:
            # contiguous with a tail, aligned 2-D rows with a tail, then
            # layouts the vector path declines (odd offset, 1001 pitch)
            for view, vd in ((a, ad), (b[:, :1000], bd[:, :1000]), (b[:, ::-1], bd[:, ::-1]),
                             (b[:, 999::-1], bd[:, 999::-1]), (a[:, 1:], ad[:, 1:]), (a[:7, :992], ad[:7, :992])):
                want = O.cast(view, to)
                got = vd.astype(to).get()
                assert got.tobytes() == want.tobytes(), ("cast", np.dtype(src).name, np.dtype(to).name, view.shape)
    # large: inputs beyond the Infinity Cache take nontemporal loads
    big = rng.random((8192, 8192)) < 0.5
    xb = rng.standard_normal((8192, 8192)).astype(np.float32)
    yb = rng.standard_normal((8192, 1)).astype(np.float32)
    got = ap.where(ap.asarray(big), ap.asarray(xb), ap.asarray(yb)).get()
    assert got.tobytes() == O.where(big, xb, yb).tobytes()
    del big, yb, got
    xd = xb.astype(np.float64) * 1e35
    assert ap.asarray(xd).astype(np.float32).get().tobytes() == O.cast(xd, np.float32).tobytes()
