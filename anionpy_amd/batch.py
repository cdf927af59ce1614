"""Batched execution of many small ufunc / reduce calls (ionp_batch_* in
include/ionp_hip.h).  No counterpart in the reference, which runs one call at
a time; every mode returns the same bytes as the single-call entry points.

    calls = [binary_call("add", a, b), reduce_call(x, axis=1), ...]
    with Batch(calls) as bt:
        bt.run("persistent")     # or "graph", "per_call"
    calls[0].out.get()
"""
import ctypes

import numpy as np

from . import _lib
from . import _CODE, _NP_DTYPES, _call, _dtype, asarray, empty_strided

_OPS = {"add": _lib.ADD, "subtract": _lib.SUBTRACT, "multiply": _lib.MULTIPLY, "true_divide": _lib.TRUE_DIVIDE,
        "divide": _lib.TRUE_DIVIDE, "maximum": _lib.MAXIMUM, "minimum": _lib.MINIMUM, "less": _lib.LESS,
        "less_equal": _lib.LESS_EQUAL, "greater": _lib.GREATER, "greater_equal": _lib.GREATER_EQUAL,
        "equal": _lib.EQUAL, "not_equal": _lib.NOT_EQUAL}
MODES = {"persistent": _lib.BATCH_PERSISTENT, "graph": _lib.BATCH_GRAPH, "per_call": _lib.BATCH_PER_CALL}


class CallSpec:
    """One call: the ionp_call record plus the arrays it references."""

    def __init__(self, kind, op, inputs, out, naxes=0, axes=(), keepdims=False):
        self.kind = kind
        self.op = op
        self.inputs = list(inputs)
        self.out = out
        c = _lib.Call()
        c.kind = kind
        c.op = op
        c.naxes = naxes
        c.keepdims = int(bool(keepdims))
        for i, a in enumerate(axes):
            c.axes[i] = a
        for i, x in enumerate(self.inputs):
            c.inp[i] = x._desc()
        c.out = out._desc()
        self.rec = c

    def features(self):
        f = (ctypes.c_float * _lib.NFEATURES)()
        _call(_lib.load().ionp_call_features, ctypes.byref(self.rec), f)
        return np.frombuffer(f, dtype=np.float32).copy()

    @property
    def nbytes(self):
        return sum(x.nbytes for x in self.inputs) + self.out.nbytes


def _template_out(t):
    return empty_strided(tuple(t.shape[: t.ndim]), tuple(t.strides[: t.ndim]), _NP_DTYPES[t.dtype])


def binary_call(op, a, b, out=None):
    a, b = asarray(a), asarray(b)
    code = _OPS[op]
    if out is None:
        t = _lib.ArrayDesc()
        da, db = a._desc(), b._desc()
        _call(_lib.load().ionp_plan_binary, code, ctypes.byref(da), ctypes.byref(db), ctypes.byref(t))
        out = _template_out(t)
    return CallSpec(_lib.CALL_BINARY, code, [a, b], out)


def where_call(cond, x, y, out=None):
    c, x, y = asarray(cond), asarray(x), asarray(y)
    if out is None:
        t = _lib.ArrayDesc()
        dc, dx, dy = c._desc(), x._desc(), y._desc()
        _call(_lib.load().ionp_plan_where, ctypes.byref(dc), ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(t))
        out = _template_out(t)
    return CallSpec(_lib.CALL_WHERE, 0, [c, x, y], out)


def cast_call(x, dtype, out=None):
    x = asarray(x)
    if out is None:
        t = _lib.ArrayDesc()
        dx = x._desc()
        _call(_lib.load().ionp_plan_cast, ctypes.byref(dx), _CODE[_dtype(dtype)], ctypes.byref(t))
        out = _template_out(t)
    return CallSpec(_lib.CALL_CAST, 0, [x], out)


def reduce_call(x, axis=0, keepdims=False, dtype=None, op="add", out=None):
    x = asarray(x)
    if axis is None:
        axes, naxes = (), -1
    else:
        axes = (axis,) if isinstance(axis, int) else tuple(axis)
        naxes = len(axes)
    code = _OPS[op]
    if out is None:
        t = _lib.ArrayDesc()
        dx = x._desc()
        arr = (ctypes.c_int32 * max(len(axes), 1))(*axes)
        dt = -1 if dtype is None else _CODE[_dtype(dtype)]
        _call(_lib.load().ionp_plan_reduce, code, ctypes.byref(dx), arr if naxes >= 0 else None, naxes,
              int(bool(keepdims)), dt, ctypes.byref(t))
        out = _template_out(t)
    return CallSpec(_lib.CALL_REDUCE, code, [x], out, naxes=naxes, axes=axes, keepdims=keepdims)


def call_array(specs):
    arr = (_lib.Call * max(len(specs), 1))()
    for i, s in enumerate(specs):
        arr[i] = s.rec
    return arr


# tile schedules the batch autotuner tries (entry 0: the library's defaults)
BATCH_GRID = [(0, 0, 0)] + [(e, i, c) for e in (8192, 16384, 32768, 65536) for i in (4096, 8192, 16384)
                            for c in (1, 2, 4, 8)]


def autotune(specs, grid=BATCH_GRID, reps=5):
    """Persistent-run time (us, fastest of `reps`) of a batch of `specs`
    under each (ew_tile, int_seg, claim) schedule of `grid`; negative:
    unusable.  Outputs are bit-identical whatever the schedule."""
    cands = (_lib.BatchConfig * len(grid))()
    for j, (e, i, c) in enumerate(grid):
        cands[j].ew_tile, cands[j].int_seg, cands[j].claim = e, i, c
    t = (ctypes.c_float * len(grid))()
    arr = call_array(specs)
    _call(_lib.load().ionp_batch_autotune, arr, len(specs), cands, len(grid), reps, t, _lib.stream_handle())
    return np.frombuffer(t, dtype=np.float32).copy()


class Batch:
    """A planned batch; `run(mode)` executes it on torch's current stream."""

    def __init__(self, specs, ew_tile=0, int_seg=0, claim=0):
        """ew_tile / int_seg / claim: tile sizes and tiles per claim of the
        persistent kernel (0: the library's defaults; see ionp_batch_config)."""
        self.specs = list(specs)
        self._arr = call_array(self.specs)
        self._h = ctypes.c_void_p()
        cfg = _lib.BatchConfig()
        cfg.ew_tile, cfg.int_seg, cfg.claim = int(ew_tile), int(int_seg), int(claim)
        _call(_lib.load().ionp_batch_create_ex, self._arr, len(self.specs), ctypes.byref(cfg), _lib.stream_handle(),
              ctypes.byref(self._h))

    @classmethod
    def tuned(cls, specs, grid=BATCH_GRID, reps=5):
        """A batch built with the fastest tile schedule of `grid` for these
        calls (the autotuner's timings in .tune_times, the schedule in
        .schedule)."""
        t = autotune(specs, grid, reps)
        j = int(np.argmin(np.where(t > 0, t, np.inf)))
        bt = cls(specs, *grid[j])
        bt.tune_times, bt.schedule = t, grid[j]
        return bt

    def run(self, mode="persistent"):
        _call(_lib.load().ionp_batch_run, self._h, MODES[mode], _lib.stream_handle())

    def info(self):
        nt, nb = ctypes.c_int64(), ctypes.c_int64()
        _call(_lib.load().ionp_batch_info, self._h, ctypes.byref(nt), ctypes.byref(nb))
        return {"tiles": nt.value, "bytes": nb.value, "calls": len(self.specs)}

    def close(self):
        if self._h:
            _lib.load().ionp_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
