"""Golden-vector cases (tests/golden): loading and operand reconstruction.

A case input is either a view into a stored buffer (offset/shape/strides in
elements) or a seeded generator spec; expected results are values, a
SHA-256 of the C-order bytes, or an exception name and message.
"""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

_cache = {}


def load():
    if "cases" not in _cache:
        with open(os.path.join(GOLDEN, "cases.json")) as f:
            _cache["cases"] = json.load(f)
        _cache["npz"] = np.load(os.path.join(GOLDEN, "fixtures.npz"))
    return _cache["cases"], _cache["npz"]


def big():
    p = os.path.join(GOLDEN, "big.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def py_scalar(spec):
    """A Python bool / int / float operand recorded as its type and exact value."""
    t, v = spec["py"], spec["value"]
    if t == "bool":
        return v == "True"
    if t == "int":
        return int(v)
    return float.fromhex(v)


def host_input(spec, npz):
    """The operand as a host numpy view (strides preserved), or the Python
    scalar itself."""
    if "py" in spec:
        return py_scalar(spec)
    if "gen" in spec:  # shift + scale * uniform[-1, 1) in float64, cast, reshaped, stepped
        x = np.random.default_rng(spec["seed"]).uniform(-1, 1, spec["n"])
        if "scale" in spec:
            x = spec.get("shift", 0.0) + spec["scale"] * x
        x = x.astype(spec["dtype"])
        if "shape" in spec:
            return x.reshape(spec["shape"])
        return x[:: spec["step"]]
    buf = npz[spec["buf"]]
    isz = buf.itemsize
    if buf.size == 0 or 0 in spec["shape"]:
        return np.zeros(spec["shape"], dtype=buf.dtype)
    return np.lib.stride_tricks.as_strided(buf[spec["offset"]:], shape=spec["shape"],
                                           strides=[s * isz for s in spec["strides"]])


def args_axis(args):
    ax = args.get("axis", 0)
    if isinstance(ax, list):
        ax = tuple(ax)
    return ax


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_result(case, got, npz):
    """Compare a host result with the expectation; returns '' or a reason."""
    exp = case["expected"]
    got = np.asarray(got)
    if got.dtype.name != exp["dtype"]:
        return f"dtype {got.dtype.name} != {exp['dtype']}"
    if list(got.shape) != exp["shape"]:
        return f"shape {got.shape} != {exp['shape']}"
    if exp.get("check_strides", True) and got.ndim and got.size > 1:
        st = [s // got.itemsize for s in got.strides]
        want = exp["strides"]
        if any(n > 1 and a != b for n, a, b in zip(got.shape, st, want)):
            return f"strides {st} != {want}"
    if "sha256" in exp:
        return "" if sha(got) == exp["sha256"] else "value digest differs"
    want = npz[exp["buf"]].reshape(exp["shape"])
    if np.ascontiguousarray(got).tobytes() != want.tobytes():
        diff = np.flatnonzero(np.ascontiguousarray(got).reshape(-1).view(np.uint8) != want.reshape(-1).view(np.uint8))
        return f"value bytes differ at byte {diff[:4]} (got {got.reshape(-1)[:3]} want {want.reshape(-1)[:3]})"
    return ""
