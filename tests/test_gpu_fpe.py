"""np.seterr / np.errstate on the GPU path: the device flag word reaches the
warning machinery with NumPy's messages and modes (flags themselves are
checked case by case against NumPy in the golden "fpe" kind)."""
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_default_modes_warn_like_numpy(ap):
    a = ap.asarray(np.array([1.0, 0.0, 2.0]))
    b = ap.asarray(np.array([0.0, 0.0, 1.0]))
    with ap.errstate(divide="warn", over="warn", under="ignore", invalid="warn"):
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            ap.divide(a, b)
    assert [str(x.message) for x in w] == ["divide by zero encountered in divide",
                                           "invalid value encountered in divide"]
    assert all(x.category is RuntimeWarning for x in w)


def test_raise_call_log_modes(ap):
    x = ap.asarray(np.array([1e308, 1.0]))
    with ap.errstate(over="raise"):
        with pytest.raises(FloatingPointError, match="overflow encountered in multiply"):
            ap.multiply(x, x)
    got = []
    with ap.errstate(all="call", call=lambda t, f: got.append((t, f))):
        ap.multiply(x, x)
        ap.multiply(ap.asarray(np.array([1e-300])), ap.asarray(np.array([1e-300])))
    assert got == [("overflow", 2), ("underflow", 4)]

    class Log:
        def __init__(self):
            self.lines = []

        def write(self, s):
            self.lines.append(s)

    log = Log()
    with ap.errstate(all="log", call=log):
        ap.asarray(np.array([np.nan, 3e9])).astype(np.int32)
    assert log.lines == ["Warning: invalid value encountered in cast\n"]


def test_ignore_and_flag_word_cleared(ap):
    x = ap.asarray(np.array([1e308]))
    with ap.errstate(all="ignore"):
        ap.multiply(x, x)  # not tracked: nothing is left behind
    got = []
    with ap.errstate(all="call", call=lambda t, f: got.append(f)):
        ap.add(x, x)
        ap.add(ap.asarray(np.array([1.0])), ap.asarray(np.array([2.0])))
    assert got == [2]


def test_seterr_roundtrip(ap):
    old = ap.seterr(all="raise", under="ignore")
    try:
        assert ap.geterr() == {"divide": "raise", "over": "raise", "under": "ignore", "invalid": "raise"}
        with pytest.raises(ValueError):
            ap.seterr(over="bogus")
    finally:
        ap.seterr(**old)
