"""CPU oracle for the ufunc / reduction hot path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package; the product package ``anionpy_amd`` never does.

The arithmetic lives in ``ionp_oracle.c`` (a scalar C restatement of the
reference's algorithm, see the citations there); this module only marshals
NumPy arrays (any strides, including negative and zero) into its descriptors
and allocates outputs with NumPy's K-order layout.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MAXDIM = 16
DTYPES = [np.bool_, np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32,
          np.int64, np.uint64, np.float32, np.float64, np.float16]
_CODE = {np.dtype(d): i for i, d in enumerate(DTYPES)}

OPS = ["add", "subtract", "multiply", "true_divide", "maximum", "minimum",
       "less", "less_equal", "greater", "greater_equal", "equal", "not_equal"]
OP = {n: i for i, n in enumerate(OPS)}

BUFSIZE = 8192  # np.getbufsize() of NumPy 2.2.6


class _Desc(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("ndim", ctypes.c_int32),
                ("shape", ctypes.c_int64 * MAXDIM), ("strides", ctypes.c_int64 * MAXDIM),
                ("data", ctypes.c_void_p)]


class OracleError(Exception):
    def __init__(self, code):
        super().__init__(f"oracle status {code}")
        self.code = code


def build():
    """Compile liboracle.so from ionp_oracle.c (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        src = os.path.join(_HERE, "ionp_oracle.c")
        if not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(_Desc)
        for name, args in [("ora_binary", [ctypes.c_int, P, P, P]),
                           ("ora_where", [P, P, P, P]),
                           ("ora_cast", [P, P]),
                           ("ora_reduce", [ctypes.c_int, P, ctypes.c_char_p, P, ctypes.c_int64, ctypes.c_void_p]),
                           ("ora_reduce_where", [ctypes.c_int, P, ctypes.c_char_p, P, P, ctypes.c_int64,
                                                 ctypes.c_void_p]),
                           ("ora_binary_omp", [ctypes.c_int, P, P, P, ctypes.POINTER(ctypes.c_int)]),
                           ("ora_binary_fpe", [ctypes.c_int, P, P, P, ctypes.POINTER(ctypes.c_int)]),
                           ("ora_cast_fpe", [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
                           ("ora_reduceat", [ctypes.c_int, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                             ctypes.c_int64, P]),
                           ("ora_argext", [ctypes.c_int, P, ctypes.c_int, P]),
                           ("ora_accumulate", [ctypes.c_int, P, ctypes.c_int, P])]:
            getattr(L, name).argtypes = args
            getattr(L, name).restype = ctypes.c_int
        L.ora_result_type.argtypes = [ctypes.c_int, ctypes.c_int]
        L.ora_result_type.restype = ctypes.c_int
        L.ora_binary_types.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.ora_binary_types.restype = ctypes.c_int
        L.ora_reduce_dtype.argtypes = [ctypes.c_int, ctypes.c_int]
        L.ora_reduce_dtype.restype = ctypes.c_int
        L.ora_korder_strides.argtypes = [ctypes.c_int, ctypes.POINTER(P), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.ora_korder_strides.restype = ctypes.c_int
        L.ora_pairwise_f64.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.ora_pairwise_f64.restype = ctypes.c_double
        L.ora_pairwise_f32.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.ora_pairwise_f32.restype = ctypes.c_float
        L.ora_axis_order.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]
        L.ora_axis_order.restype = None
        _lib = L
    return _lib


def _desc(a):
    a = np.asarray(a)
    d = _Desc()
    d.dtype = _CODE[a.dtype]
    d.ndim = a.ndim
    for i in range(a.ndim):
        d.shape[i] = a.shape[i]
        assert a.strides[i] % a.itemsize == 0
        d.strides[i] = a.strides[i] // a.itemsize
    d.data = a.__array_interface__["data"][0]
    return d


def _check(rc):
    if rc:
        raise OracleError(rc)


def _empty_strided(shape, strides_elems, dtype):
    dtype = np.dtype(dtype)
    span = 1 + sum((s - 1) * abs(st) for s, st in zip(shape, strides_elems)) if all(shape) else 0
    base = np.empty(max(span, 1), dtype=dtype)
    return np.lib.stride_tricks.as_strided(base, shape=shape,
                                           strides=[st * dtype.itemsize for st in strides_elems])


def korder_strides(arrays, shape):
    descs = [_desc(a) for a in arrays]
    arr = (ctypes.POINTER(_Desc) * len(descs))(*[ctypes.pointer(d) for d in descs])
    shp = (ctypes.c_int64 * MAXDIM)(*shape)
    out = (ctypes.c_int64 * MAXDIM)()
    _check(lib().ora_korder_strides(len(descs), arr, len(shape), shp, out))
    return [out[i] for i in range(len(shape))]


def result_type(a, b):
    return np.dtype(DTYPES[lib().ora_result_type(_CODE[np.dtype(a)], _CODE[np.dtype(b)])])


def binary_types(op, a, b):
    lo, ou = ctypes.c_int(), ctypes.c_int()
    _check(lib().ora_binary_types(OP[op], _CODE[np.dtype(a)], _CODE[np.dtype(b)],
                                  ctypes.byref(lo), ctypes.byref(ou)))
    return np.dtype(DTYPES[lo.value]), np.dtype(DTYPES[ou.value])


def broadcast_shape(*shapes):
    nd = max(len(s) for s in shapes)
    out = []
    for i in range(nd):
        dims = [s[i - (nd - len(s))] if i >= nd - len(s) else 1 for s in shapes]
        big = [d for d in dims if d != 1]
        if big and any(d != big[0] for d in big):
            raise OracleError(1)
        out.append(big[0] if big else 1)
    return tuple(out)


def binary(op, a, b):
    """np.<op>(a, b) with NumPy's K-order output layout."""
    a = np.asarray(a)
    b = np.asarray(b)
    _, odt = binary_types(op, a.dtype, b.dtype)
    shape = broadcast_shape(a.shape, b.shape)
    out = _empty_strided(shape, korder_strides([a, b], shape), odt)
    da, db, do = _desc(a), _desc(b), _desc(out)
    _check(lib().ora_binary(OP[op], ctypes.byref(da), ctypes.byref(db), ctypes.byref(do)))
    return out


def binary_omp(op, a, b):
    """binary() on all host cores (OpenMP, rows split per thread): the CPU
    baseline's multi-thread leg.  Returns (out, threads used)."""
    a = np.asarray(a)
    b = np.asarray(b)
    _, odt = binary_types(op, a.dtype, b.dtype)
    shape = broadcast_shape(a.shape, b.shape)
    out = _empty_strided(shape, korder_strides([a, b], shape), odt)
    da, db, do = _desc(a), _desc(b), _desc(out)
    nt = ctypes.c_int()
    _check(lib().ora_binary_omp(OP[op], ctypes.byref(da), ctypes.byref(db), ctypes.byref(do), ctypes.byref(nt)))
    return out, nt.value


def binary_fpe(op, a, b):
    """IEEE flags (1 divide, 2 over, 4 under, 8 invalid) that np.<op>(a, b)
    raises, for op in add/subtract/multiply/true_divide."""
    a = np.asarray(a)
    b = np.asarray(b)
    _, odt = binary_types(op, a.dtype, b.dtype)
    shape = broadcast_shape(a.shape, b.shape)
    out = _empty_strided(shape, korder_strides([a, b], shape), odt)
    fl = ctypes.c_int()
    _check(lib().ora_binary_fpe(OP[op], ctypes.byref(_desc(a)), ctypes.byref(_desc(b)), ctypes.byref(_desc(out)),
                                ctypes.byref(fl)))
    return fl.value


def cast_fpe(x, dtype):
    """IEEE flags that x.astype(dtype) raises."""
    x = np.asarray(x)
    fl = ctypes.c_int()
    _check(lib().ora_cast_fpe(ctypes.byref(_desc(x)), _CODE[np.dtype(dtype)], ctypes.byref(fl)))
    return fl.value


def where(c, x, y):
    c, x, y = np.asarray(c), np.asarray(x), np.asarray(y)
    odt = result_type(x.dtype, y.dtype)
    shape = broadcast_shape(c.shape, x.shape, y.shape)
    out = _empty_strided(shape, korder_strides([c, x, y], shape), odt)
    dc, dx, dy, do = _desc(c), _desc(x), _desc(y), _desc(out)
    _check(lib().ora_where(ctypes.byref(dc), ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(do)))
    return out


def cast(x, dtype):
    """x.astype(dtype) (order='K')."""
    x = np.asarray(x)
    out = _empty_strided(x.shape, korder_strides([x], x.shape), dtype)
    dx, do = _desc(x), _desc(out)
    _check(lib().ora_cast(ctypes.byref(dx), ctypes.byref(do)))
    return out


def _norm_axes(axis, ndim):
    if axis is None:
        return list(range(ndim))
    if isinstance(axis, int):
        axis = (axis,)
    out = []
    for a in axis:
        n = a + max(ndim, 1) if a < 0 else a
        if n < 0 or n >= max(ndim, 1):
            raise OracleError(6)
        out.append(n)
    if len(set(out)) != len(out):
        raise OracleError(4)
    return [] if ndim == 0 else out


def reduce(op, x, axis=0, bufsize=BUFSIZE, dtype=None, initial=None, where=None):
    """np.<op>.reduce(x, axis, dtype, initial, where) for op in
    add/multiply/maximum/minimum.  `where`: a bool mask broadcastable to x
    (ora_reduce_where), else the unmasked fold (ora_reduce)."""
    x = np.asarray(x)
    axes = _norm_axes(axis, x.ndim)
    mask = bytes(1 if k in axes else 0 for k in range(x.ndim)) or b"\0"
    odt = DTYPES[lib().ora_reduce_dtype(OP[op], _CODE[x.dtype])] if dtype is None else np.dtype(dtype)
    if odt in (np.float32, np.float64) and x.dtype != odt and (x.dtype == np.float16 or op in ("maximum", "minimum")):
        # NumPy's buffered cast ahead of a float32 / float64 loop (ora_reduce
        # reads float inputs as they are): the loop sees the cast values, the
        # iterator's axis order unchanged
        x = x.astype(odt, order="K")
    kept = [k for k in range(x.ndim) if k not in axes]
    shape = tuple(x.shape[k] for k in kept)
    wm = None
    if where is not None:
        wm = np.broadcast_to(np.asarray(where), x.shape)
        if wm.dtype != np.bool_:
            raise OracleError(3)
    # K-order of the kept axes, following the input's iterator order (with
    # a mask: the order over input and mask)
    order = []
    if x.ndim and wm is None:
        perm = (ctypes.c_int * MAXDIM)()
        shp = (ctypes.c_int64 * MAXDIM)(*x.shape)
        st = (ctypes.c_int64 * MAXDIM)(*[s // x.itemsize for s in x.strides])
        lib().ora_axis_order(x.ndim, shp, st, perm)
        order = [perm[i] for i in range(x.ndim) if perm[i] in kept]
    elif x.ndim:
        full = korder_strides([x, wm], x.shape)
        order = sorted(kept, key=lambda a: (full[a], -a))
    ostr = [0] * len(kept)
    acc = 1
    for a in order:
        ostr[kept.index(a)] = acc
        acc *= max(x.shape[a], 1)
    out = _empty_strided(shape, ostr, odt)
    dx, do = _desc(x), _desc(out)
    init = None if initial is None else np.array(initial, dtype=odt).reshape(1)
    ip = None if init is None else init.ctypes.data
    if wm is None:
        _check(lib().ora_reduce(OP[op], ctypes.byref(dx), mask, ctypes.byref(do), bufsize, ip))
    else:
        dm = _desc(wm)
        _check(lib().ora_reduce_where(OP[op], ctypes.byref(dx), mask, ctypes.byref(dm), ctypes.byref(do),
                                      bufsize, ip))
    return out[()] if out.ndim == 0 else out


def reduceat(op, x, indices, axis=0, dtype=None):
    """np.<op>.reduceat(x, indices, axis, dtype) for op in
    add/multiply/maximum/minimum, with NumPy's output layout."""
    x = np.asarray(x)
    if x.ndim == 0:
        raise OracleError(3)
    ax = axis + x.ndim if axis < 0 else axis
    if not 0 <= ax < x.ndim:
        raise OracleError(6)
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int64).reshape(-1))
    n = x.shape[ax]
    for i in idx:  # checked in call order (validate_reduceat_indices, ufunc.rs:8426)
        if i < 0 or i >= n:
            raise OracleError(2)
    odt = DTYPES[lib().ora_reduce_dtype(OP[op], _CODE[x.dtype])] if dtype is None else np.dtype(dtype)
    shape = list(x.shape)
    shape[ax] = len(idx)
    # output layout: NumPy's iterator over (input without `axis`, indices
    # along `axis`) in K order -- `axis` never compares with the others
    xs = list(x.strides)
    xs[ax] = 0
    iv = np.lib.stride_tricks.as_strided(x, shape=shape, strides=xs)
    ind = np.zeros(max(len(idx), 1), dtype=np.int64)
    ist = [0] * x.ndim
    ist[ax] = 8
    jv = np.lib.stride_tricks.as_strided(ind, shape=shape, strides=ist)
    out = _empty_strided(tuple(shape), korder_strides([iv, jv], shape), odt)
    dx, do = _desc(x), _desc(out)
    pi = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    _check(lib().ora_reduceat(OP[op], ctypes.byref(dx), ax, pi, len(idx), ctypes.byref(do)))
    return out


def accumulate(op, x, axis=0, dtype=None):
    """np.<op>.accumulate(x, axis, dtype) for op in add/maximum/minimum, with
    the output laid out in the input's K order."""
    x = np.asarray(x)
    if x.ndim == 0:
        raise OracleError(6)
    ax = axis + x.ndim if axis < 0 else axis
    odt = DTYPES[lib().ora_reduce_dtype(OP[op], _CODE[x.dtype])] if dtype is None else np.dtype(dtype)
    out = _empty_strided(x.shape, korder_strides([x], x.shape), odt)
    dx, do = _desc(x), _desc(out)
    _check(lib().ora_accumulate(OP[op], ctypes.byref(dx), ax, ctypes.byref(do)))
    return out


def argext(ismax, x, axis=None):
    x = np.asarray(x)
    if axis is None:
        out = np.empty((), dtype=np.int64)
        ax = -1
    else:
        ax = _norm_axes(axis, x.ndim)[0]
        out = np.empty(tuple(s for k, s in enumerate(x.shape) if k != ax), dtype=np.int64)
    dx, do = _desc(x), _desc(out)
    _check(lib().ora_argext(1 if ismax else 0, ctypes.byref(dx), ax, ctypes.byref(do)))
    return out[()] if out.ndim == 0 else out


def argmax(x, axis=None):
    return argext(True, x, axis)


def argmin(x, axis=None):
    return argext(False, x, axis)


def pairwise(x):
    """NumPy's pairwise tree over one contiguous run (no buffer windows)."""
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return np.float64(lib().ora_pairwise_f64(x.ctypes.data, x.size))
    return np.float32(lib().ora_pairwise_f32(x.ctypes.data, x.size))
