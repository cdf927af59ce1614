"""Seeded generators of call layouts.

uneven_batch(seed, n)  - the uneven batch workload: n calls whose total size
                         is log-uniform in [16, 2^20] elements, rank 1-4,
                         each axis made size-1 in one operand with p = 0.25
                         (broadcast), dtypes f32/f64/i64 per operand, ops
                         add / multiply / less / where / add.reduce (random
                         axis or None).
tuning_layouts(seed, n) - layouts for the autotuner and the cost model:
                         elementwise and reduce calls over random shapes,
                         strides (transposes, steps, negative steps,
                         broadcasts), dtypes and op classes, carved out of
                         one device pool (values are irrelevant to timing).
"""
import math

import numpy as np
import torch

from . import _device, ndarray
from .batch import binary_call, cast_call, reduce_call, where_call

_BATCH_DTYPES = (np.float32, np.float64, np.int64)
_BATCH_OPS = ("add", "multiply", "less", "where", "add.reduce")


def _split_shape(rng, total, rank):
    """`rank` factors whose product is close to `total` (each >= 1)."""
    logs = rng.dirichlet(np.ones(rank)) * math.log(max(total, 1))
    shape = [max(1, int(round(math.exp(v)))) for v in logs]
    return tuple(shape)


def _device_array(shape, dtype, gen):
    """Fresh C-contiguous device array with seeded values."""
    n = int(np.prod(shape)) if shape else 1
    dt = np.dtype(dtype)
    if dt.kind == "f":
        t = torch.rand(max(n, 1), generator=gen, device=_device(), dtype=torch.float64) * 2 - 1
        t = t.to(torch.float32 if dt == np.float32 else torch.float64)
    elif dt.kind == "b":
        t = torch.rand(max(n, 1), generator=gen, device=_device()) < 0.5
    else:
        t = torch.randint(-1000, 1000, (max(n, 1),), generator=gen, device=_device(), dtype=torch.int64)
        t = t.to({1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[dt.itemsize])
    buf = t.view(torch.uint8).reshape(-1)
    st, acc = [0] * len(shape), 1
    for i in range(len(shape) - 1, -1, -1):
        st[i] = acc
        acc *= shape[i]
    return ndarray(buf, dt, shape, st, 0)


def uneven_batch(seed=5, n=10000, max_log2=20, min_elems=16):
    """The uneven batch: list of CallSpec over fresh device arrays."""
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=_device())
    gen.manual_seed(seed)
    specs = []
    for _ in range(n):
        total = int(math.exp(rng.uniform(math.log(min_elems), math.log(2 ** max_log2))))
        rank = int(rng.integers(1, 5))
        shape = _split_shape(rng, total, rank)
        op = _BATCH_OPS[int(rng.integers(0, len(_BATCH_OPS)))]
        dts = [_BATCH_DTYPES[int(rng.integers(0, 3))] for _ in range(3)]
        if op == "add.reduce":
            x = _device_array(shape, dts[0], gen)
            axes = [None] + list(range(rank))
            specs.append(reduce_call(x, axis=axes[int(rng.integers(0, len(axes)))]))
            continue
        # broadcast: every axis is size-1 in one of the operands with p = 0.25
        nin = 3 if op == "where" else 2
        shapes = [list(shape) for _ in range(nin)]
        for ax in range(rank):
            if rng.random() < 0.25:
                shapes[int(rng.integers(0, nin))][ax] = 1
        if op == "where":
            c = _device_array(tuple(shapes[0]), np.bool_, gen)
            x = _device_array(tuple(shapes[1]), dts[1], gen)
            y = _device_array(tuple(shapes[2]), dts[2], gen)
            specs.append(where_call(c, x, y))
        else:
            a = _device_array(tuple(shapes[0]), dts[0], gen)
            b = _device_array(tuple(shapes[1]), dts[1], gen)
            specs.append(binary_call(op, a, b))
    return specs


# ---- tuning layouts ---------------------------------------------------------
class Pool:
    """One device allocation that views are carved from."""

    def __init__(self, nbytes=1 << 29):
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=_device())
        self.buf.zero_()
        self.nbytes = nbytes

    def view(self, rng, shape, dtype, kind):
        """A view of `shape` with a layout of the given kind."""
        dt = np.dtype(dtype)
        isz = dt.itemsize
        nd = len(shape)
        if kind == "transposed" and nd >= 2:
            perm = list(rng.permutation(nd))
            base = [shape[p] for p in perm]
            st_base = _c_strides(base)
            st = [0] * nd
            for i, p in enumerate(perm):
                st[p] = st_base[i]
        elif kind == "stepped":
            steps = [int(rng.integers(1, 3)) for _ in range(nd)]
            base = [s * k for s, k in zip(shape, steps)]
            st = [a * k for a, k in zip(_c_strides(base), steps)]
        elif kind == "reversed":
            st = _c_strides(shape)
            ax = int(rng.integers(0, nd))
            st[ax] = -st[ax]
        else:
            st = _c_strides(shape)
        lo = sum(min(0, (n - 1) * s) for n, s in zip(shape, st))
        hi = sum(max(0, (n - 1) * s) for n, s in zip(shape, st))
        span = hi - lo + 1
        if span * isz > self.nbytes:
            raise ValueError("layout larger than the pool")
        limit = self.nbytes // isz - span
        start = int(rng.integers(0, max(limit, 1))) // 64 * 64 if limit > 0 else 0
        return ndarray(self.buf, dt, shape, st, start - lo)


def _c_strides(shape):
    st, acc = [0] * len(shape), 1
    for i in range(len(shape) - 1, -1, -1):
        st[i] = acc
        acc *= shape[i]
    return st


_TUNE_DTYPES = (np.float32, np.float64, np.int64, np.int32, np.uint8)
_KINDS = ("contiguous", "contiguous", "transposed", "stepped", "reversed")
CLASSES = ("arith", "compare", "where", "cast", "reduce")


def tuning_layouts(seed, n, pool, min_log2=10, max_log2=22):
    """n random calls for timing sweeps (see module docstring)."""
    rng = np.random.default_rng(seed)
    specs = []
    while len(specs) < n:
        total = int(2 ** rng.uniform(min_log2, max_log2))
        rank = int(rng.integers(1, 5))
        shape = _split_shape(rng, total, rank)
        cls = int(rng.integers(0, 5))
        dt = _TUNE_DTYPES[int(rng.integers(0, len(_TUNE_DTYPES)))]
        kind = _KINDS[int(rng.integers(0, len(_KINDS)))]
        try:
            if cls <= 1:  # arithmetic / comparison, with broadcasting
                b_shape = [s if rng.random() > 0.3 else 1 for s in shape]
                a = pool.view(rng, shape, dt, kind)
                b = pool.view(rng, tuple(b_shape), dt, "contiguous")
                op = ("add", "multiply", "maximum")[int(rng.integers(0, 3))] if cls == 0 else "less"
                spec = binary_call(op, a, b, out=_out_for(pool, rng, shape, np.bool_ if cls else dt))
            elif cls == 2:
                c = pool.view(rng, shape, np.bool_, "contiguous")
                x = pool.view(rng, shape, dt, kind)
                y = pool.view(rng, tuple(1 for _ in shape), dt, "contiguous")
                spec = where_call(c, x, y, out=_out_for(pool, rng, shape, dt))
            elif cls == 3:
                x = pool.view(rng, shape, dt, kind)
                to = _TUNE_DTYPES[int(rng.integers(0, len(_TUNE_DTYPES)))]
                spec = cast_call(x, to, out=_out_for(pool, rng, shape, to))
            else:
                x = pool.view(rng, shape, dt if dt != np.uint8 else np.float64, kind)
                axes = [None] + list(range(rank))
                spec = reduce_call(x, axis=axes[int(rng.integers(0, len(axes)))])
        except (ValueError, TypeError, NotImplementedError):
            continue
        spec.tag = (CLASSES[cls], kind, total)
        specs.append(spec)
    return specs


def _out_for(pool, rng, shape, dtype):
    return pool.view(rng, shape, dtype, "contiguous")
