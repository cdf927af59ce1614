"""anionpy_amd -- MI355X (gfx950) backend for anionpy's ufunc and reduction engine.

NumPy-named API over device arrays, mirroring the reference's Python surface
(anionpy/__init__.py re-exports of the ionp-py Ufunc objects,
ionp-py/src/lib.rs:8697 `Ufunc`, ionp-py/src/reductions.rs sum/max/argmax,
where).  Every call runs on the GPU through the C ABI in include/ionp_hip.h;
there is no CPU fallback: without the built library, import fails.

    import anionpy_amd as ap
    A = ap.asarray(np_array)            # host -> HBM
    C = ap.add(A, b)                     # broadcast, K-order output
    s = ap.add.reduce(C, axis=0)         # NumPy's pairwise tree per buffer window
    C.get()                              # HBM -> host numpy (same strides)
"""
import builtins
import contextlib
import ctypes
import math
import numbers
import operator
import re
import sys
import threading
import warnings

import numpy as np
import torch

from . import _lib
from ._lib import IonpError

__all__ = ["ndarray", "asarray", "empty", "zeros", "Ufunc", "add", "subtract", "multiply", "divide",
           "true_divide", "maximum", "minimum", "less", "less_equal", "greater", "greater_equal", "equal",
           "not_equal", "where", "sum", "prod", "cumsum", "cumprod", "max", "amax", "min", "amin", "argmax", "argmin", "result_type",
           "AxisError", "UFuncTypeError", "IonpError", "setbufsize", "getbufsize", "set_config_policy",
           "get_config_policy", "prod", "cumprod", "seterr", "geterr", "seterrcall", "geterrcall", "errstate"]

_NP_DTYPES = [np.bool_, np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64,
              np.uint64, np.float32, np.float64, np.float16]
_CODE = {np.dtype(d): i for i, d in enumerate(_NP_DTYPES)}


class AxisError(np.exceptions.AxisError):
    """The reference's AxisError (ionp-py/src/errors.rs, raised by
    lib.rs:76-80 axis_error): a subclass of numpy.exceptions.AxisError (so
    also of ValueError and IndexError), built as AxisError(axis, ndim)."""


class UFuncTypeError(TypeError):
    """numpy's _UFuncNoLoopError counterpart (error.rs NoUfuncLoop)."""


def _raise(e):
    # IonpError -> Python exception type (lib.rs:65-75 to_py_err)
    code = e.code
    msg = str(e)
    if code in (1, 2, 5):
        raise ValueError(msg) from None
    if code == 3:
        raise IndexError(msg) from None
    if code == 4:
        raise TypeError(msg) from None
    if code == 6:
        raise UFuncTypeError(msg) from None
    if code == 7:
        m = re.fullmatch(r"axis (-?\d+) is out of bounds for array of dimension (\d+)", msg)
        raise (AxisError(int(m.group(1)), int(m.group(2))) if m else AxisError(msg)) from None
    if code == 21:
        raise NotImplementedError(msg) from None
    raise RuntimeError(msg) from None


def _call(fn, *args):
    code = fn(*args)
    if code != 0:
        try:
            _lib.check(code)
        except IonpError as e:
            _raise(e)


def _dtype(dt):
    dt = np.dtype(dt)
    if dt not in _CODE:
        raise TypeError(f"dtype {dt} is not supported by the MI355X backend")
    return dt


def _device():
    if not torch.cuda.is_available():
        raise RuntimeError("anionpy_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


class ndarray:
    """Strided view of device (HBM) memory: dtype, shape, element strides and
    an element offset into a byte allocation (ionp-core/src/array.rs:24-29)."""

    __slots__ = ("_buf", "dtype", "shape", "strides_el", "offset", "__weakref__")
    __array_priority__ = 100

    def __init__(self, buf, dtype, shape, strides_el, offset=0):
        self._buf = buf  # torch.uint8 CUDA tensor owning the bytes
        self.dtype = np.dtype(dtype)
        self.shape = tuple(int(s) for s in shape)
        self.strides_el = tuple(int(s) for s in strides_el)
        self.offset = int(offset)

    # ---- descriptors ---------------------------------------------------------
    def _desc(self):
        d = _lib.ArrayDesc()
        d.dtype = _CODE[self.dtype]
        d.ndim = len(self.shape)
        for i, (n, s) in enumerate(zip(self.shape, self.strides_el)):
            d.shape[i] = n
            d.strides[i] = s
        d.offset = self.offset
        d.data = self._buf.data_ptr() if self._buf is not None and self._buf.numel() else 0
        return d

    # ---- numpy-like attributes ------------------------------------------------
    @property
    def ndim(self):
        return len(self.shape)

    @property
    def size(self):
        n = 1
        for s in self.shape:
            n *= s
        return n

    @property
    def itemsize(self):
        return self.dtype.itemsize

    @property
    def strides(self):
        return tuple(s * self.itemsize for s in self.strides_el)

    @property
    def nbytes(self):
        return self.size * self.itemsize

    @property
    def T(self):
        return self.transpose()

    def __len__(self):
        if not self.shape:
            raise TypeError("len() of unsized object")
        return self.shape[0]

    def __repr__(self):
        return f"anionpy_amd.ndarray(shape={self.shape}, dtype={self.dtype}, strides={self.strides})"

    # ---- views ---------------------------------------------------------------------
    def transpose(self, *axes):
        if not axes or axes == (None,):
            axes = tuple(range(self.ndim))[::-1]
        elif len(axes) == 1 and isinstance(axes[0], (tuple, list)):
            axes = tuple(axes[0])
        axes = tuple(a % self.ndim for a in axes)
        if sorted(axes) != list(range(self.ndim)):
            raise ValueError("axes don't match array")
        return ndarray(self._buf, self.dtype, [self.shape[a] for a in axes],
                       [self.strides_el[a] for a in axes], self.offset)

    def __getitem__(self, key):
        """Basic indexing: ints, slices (any step), Ellipsis and None."""
        if not isinstance(key, tuple):
            key = (key,)
        nexp = builtins.sum(1 for k in key if k is not None and k is not Ellipsis)
        if any(k is Ellipsis for k in key):
            i = next(i for i, k in enumerate(key) if k is Ellipsis)
            key = key[:i] + (slice(None),) * (self.ndim - nexp) + key[i + 1:]
        key = key + (slice(None),) * (self.ndim - builtins.sum(1 for k in key if k is not None))
        shape, strides, off, ax = [], [], self.offset, 0
        for k in key:
            if k is None:
                shape.append(1)
                strides.append(0)
                continue
            n, st = self.shape[ax], self.strides_el[ax]
            if isinstance(k, slice):
                start, stop, step = k.indices(n)
                cnt = len(range(start, stop, step))
                off += start * st if cnt else 0
                shape.append(cnt)
                strides.append(st * step)
            elif isinstance(k, numbers.Integral):
                i = operator.index(k)
                if i < -n or i >= n:
                    raise IndexError(f"index {i} is out of bounds for axis {ax} with size {n}")
                off += (i % n) * st
            else:
                raise IndexError("only basic indexing (ints, slices, None, ...) is supported")
            ax += 1
        return ndarray(self._buf, self.dtype, shape, strides, off)

    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = list(shape)
        if -1 in shape:
            known = 1
            for s in shape:
                if s != -1:
                    known *= s
            shape[shape.index(-1)] = self.size // known if known else 0
        n = 1
        for s in shape:
            n *= s
        if n != self.size:
            raise ValueError(f"cannot reshape array of size {self.size} into shape "
                             f"({','.join(str(s) for s in shape)}{',' if len(shape) == 1 else ''})")
        src = self if self.flags_c_contiguous() else self.copy()
        st, acc = [0] * len(shape), 1
        for i in range(len(shape) - 1, -1, -1):
            st[i] = acc
            acc *= shape[i]
        return ndarray(src._buf, src.dtype, shape, st, src.offset)

    def flags_c_contiguous(self):
        acc = 1
        for n, s in zip(self.shape[::-1], self.strides_el[::-1]):
            if n != 1 and s != acc:
                return False
            acc *= n
        return True

    def flags_f_contiguous(self):
        acc = 1
        for n, s in zip(self.shape, self.strides_el):
            if n != 1 and s != acc:
                return False
            acc *= n
        return True

    def copy(self):
        return self.astype(self.dtype, order="C")

    def astype(self, dtype, order="K"):
        dtype = _dtype(dtype)
        _fpe_arm(_fpe_wanted())
        track = (self.dtype.kind == "f" or dtype == np.float16) and _fpe_wanted()  # int -> float16 can overflow
        r = self._astype(dtype, order)
        if track:
            _fpe_check("cast")
        return r

    def _astype(self, dtype, order):
        t = _lib.ArrayDesc()
        d = self._desc()
        _call(_lib.load().ionp_plan_cast, ctypes.byref(d), _CODE[dtype], ctypes.byref(t))
        strides = _c_strides(self.shape) if order == "C" else list(t.strides)[: self.ndim]
        out = empty_strided(self.shape, strides, dtype)
        o = out._desc()
        _call(_lib.load().ionp_cast, ctypes.byref(d), ctypes.byref(o), _cfg(_lib.CALL_CAST, 0, (d,), o),
              _lib.stream_handle())
        return out

    # ---- transfer ---------------------------------------------------------------------
    def get(self):
        """Copy to a host numpy array with the same shape and strides."""
        if self.size == 0:
            return np.empty(self.shape, dtype=self.dtype)
        lo, hi = _span(self.shape, self.strides_el)
        isz = self.itemsize
        nbytes = (hi - lo + 1) * isz
        start = (self.offset + lo) * isz
        _call(_lib.load().ionp_check, _lib.stream_handle())  # waits for the stream
        host = self._buf[start:start + nbytes].cpu().numpy()
        base = host.view(self.dtype)
        return np.lib.stride_tricks.as_strided(base[-lo:] if lo else base, shape=self.shape,
                                               strides=[s * isz for s in self.strides_el])

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a.astype(dtype) if dtype is not None else a

    def item(self):
        return self.get().item()

    def __float__(self):
        return float(self.item())

    def __int__(self):
        return int(self.item())

    def __bool__(self):
        return bool(self.item())

    # ---- reductions as methods (ndarray_attrs.rs mirrors do_reduce_axis) --------------
    def sum(self, axis=None, dtype=None, keepdims=False, initial=None, where=True):
        return sum(self, axis=axis, dtype=dtype, keepdims=keepdims, initial=initial, where=where)

    def cumsum(self, axis=None, dtype=None, out=None):
        return cumsum(self, axis=axis, dtype=dtype, out=out)

    def prod(self, axis=None, dtype=None, keepdims=False, initial=None, where=True):
        return prod(self, axis=axis, dtype=dtype, keepdims=keepdims, initial=initial, where=where)

    def cumprod(self, axis=None, dtype=None, out=None):
        return cumprod(self, axis=axis, dtype=dtype, out=out)

    def max(self, axis=None, keepdims=False, initial=None, where=True):
        return amax(self, axis=axis, keepdims=keepdims, initial=initial, where=where)

    def min(self, axis=None, keepdims=False, initial=None, where=True):
        return amin(self, axis=axis, keepdims=keepdims, initial=initial, where=where)

    def argmax(self, axis=None, keepdims=False):
        return argmax(self, axis=axis, keepdims=keepdims)

    def argmin(self, axis=None, keepdims=False):
        return argmin(self, axis=axis, keepdims=keepdims)

    # ---- operators -----------------------------------------------------------------------
    def __add__(self, o): return add(self, o)
    def __radd__(self, o): return add(o, self)
    def __sub__(self, o): return subtract(self, o)
    def __rsub__(self, o): return subtract(o, self)
    def __mul__(self, o): return multiply(self, o)
    def __rmul__(self, o): return multiply(o, self)
    def __truediv__(self, o): return true_divide(self, o)
    def __rtruediv__(self, o): return true_divide(o, self)
    def __lt__(self, o): return less(self, o)
    def __le__(self, o): return less_equal(self, o)
    def __gt__(self, o): return greater(self, o)
    def __ge__(self, o): return greater_equal(self, o)
    def __eq__(self, o): return equal(self, o)
    def __ne__(self, o): return not_equal(self, o)
    __hash__ = None


def _c_strides(shape):
    st, acc = [0] * len(shape), 1
    for i in range(len(shape) - 1, -1, -1):
        st[i] = acc
        acc *= builtins.max(shape[i], 1)
    return st


def _span(shape, strides):
    lo = builtins.sum(builtins.min(0, (n - 1) * s) for n, s in zip(shape, strides))
    hi = builtins.sum(builtins.max(0, (n - 1) * s) for n, s in zip(shape, strides))
    return lo, hi


def empty_strided(shape, strides_el, dtype):
    dtype = _dtype(dtype)
    shape = tuple(int(s) for s in shape)
    if any(s == 0 for s in shape):
        nel = 0
        lo = 0
    else:
        lo, hi = _span(shape, strides_el)
        nel = hi - lo + 1
    buf = torch.empty(builtins.max(nel, 1) * dtype.itemsize, dtype=torch.uint8, device=_device())
    return ndarray(buf, dtype, shape, strides_el, -lo)


def _f_strides(shape):
    st, acc = [0] * len(shape), 1
    for i in range(len(shape)):
        st[i] = acc
        acc *= builtins.max(shape[i], 1)
    return st


def _byte_range(x):
    """[first, last] byte addresses the view touches (None when empty)."""
    if x.size == 0 or x._buf is None:
        return None
    lo, hi = _span(x.shape, x.strides_el)
    base = x._buf.data_ptr()
    return base + (x.offset + lo) * x.itemsize, base + (x.offset + hi + 1) * x.itemsize - 1


def _needs_temp(out, inputs):
    """True when `out` shares memory with an input other than as the very
    same view: writing results straight into it could overwrite elements
    before they are read (NumPy then computes into a temporary)."""
    if out._buf is None or out.size == 0:
        return False
    ob = out._buf.data_ptr()
    oe = ob + out._buf.numel()
    # buffers that do not intersect cannot share an element
    if all(x._buf is None or x._buf.data_ptr() >= oe or x._buf.data_ptr() + x._buf.numel() <= ob for x in inputs):
        return False
    ro = _byte_range(out)
    if ro is None:
        return False
    for x in inputs:
        rx = _byte_range(x)
        if rx is None or rx[1] < ro[0] or ro[1] < rx[0]:
            continue
        same = (x._buf.data_ptr() == out._buf.data_ptr() and x.offset == out.offset and x.itemsize == out.itemsize
                and tuple(x.shape) == tuple(out.shape) and tuple(x.strides_el) == tuple(out.strides_el))
        if not same:
            return True
    return False


def _layout(shape, k_strides, order, inputs):
    """Output strides for order= K / C / F / A."""
    if order in (None, "K"):
        return list(k_strides)
    if order == "A":
        order = "F" if inputs and all(x.flags_f_contiguous() and not x.flags_c_contiguous() for x in inputs) else "C"
    if order == "C":
        return _c_strides(shape)
    if order == "F":
        return _f_strides(shape)
    raise ValueError(f"order must be one of 'C', 'F', 'A', or 'K' (got {order!r})")


def _copy_into(src, dst):
    """dst[...] = src (same shape; a cast kernel, exact for equal dtypes)."""
    d, o = src._desc(), dst._desc()
    _call(_lib.load().ionp_cast, ctypes.byref(d), ctypes.byref(o), None, _lib.stream_handle())


def empty(shape, dtype=np.float64):
    if isinstance(shape, numbers.Integral):
        shape = (shape,)
    return empty_strided(shape, _c_strides(shape), dtype)


def zeros(shape, dtype=np.float64):
    a = empty(shape, dtype)
    a._buf.zero_()
    return a


def asarray(x, dtype=None):
    """Host (numpy / scalar / list) -> device array.  Views keep their strides
    (including negative ones): the memory span they cover is uploaded."""
    if isinstance(x, ndarray):
        return x if dtype is None or np.dtype(dtype) == x.dtype else x.astype(dtype)
    a = np.asarray(x, dtype=dtype)
    _dtype(a.dtype)
    isz = a.itemsize
    if any(s % isz for s in a.strides) or not a.flags.aligned:
        a = np.ascontiguousarray(a)
    st = [s // isz for s in a.strides]
    if a.size == 0:
        return empty(a.shape, a.dtype)
    lo, hi = _span(a.shape, st)
    nbytes = (hi - lo + 1) * isz
    ptr = a.__array_interface__["data"][0] + lo * isz
    raw = np.ctypeslib.as_array((ctypes.c_uint8 * nbytes).from_address(ptr)).copy()
    buf = torch.from_numpy(raw).to(_device())
    return ndarray(buf, a.dtype, a.shape, st, -lo)


def setbufsize(n):
    """NumPy's iterator buffer size, which sets add.reduce's window (default 8192)."""
    old = getbufsize()
    _call(_lib.load().ionp_set_bufsize, int(n))
    return old


def getbufsize():
    return int(_lib.load().ionp_get_bufsize())


def result_type(a, b):
    out = ctypes.c_int32()
    _call(_lib.load().ionp_result_type, _CODE[_dtype(a)], _CODE[_dtype(b)], ctypes.byref(out))
    return np.dtype(_NP_DTYPES[out.value])


# ---- operands --------------------------------------------------------------------------
# NEP 50 weak Python scalars, as ionp-py marshals them (lib.rs:8238-8272
# extract_binary_pair -> scalar_against 8104, weak_target_dtype 8022): a bare
# Python bool / int / float takes the other operand's dtype within its kind;
# numpy scalars and arrays carry a dtype and are strong.
_I64, _U64, _F64, _BOOL = (np.dtype(t) for t in (np.int64, np.uint64, np.float64, np.bool_))


def _scalar_kind(x):
    """'b' / 'i' / 'f' for a weak Python bool / int / float, else None."""
    if isinstance(x, (np.generic, np.ndarray, ndarray)):
        return None
    if isinstance(x, bool):
        return "b"
    if isinstance(x, int):
        return "i"
    if isinstance(x, float):
        return "f"
    return None


def _is_weak(x):
    return _scalar_kind(x) is not None


def _weak_target(kind, od, code):
    """The dtype a weak scalar takes against an operand of dtype `od`: a
    bool takes od; an int takes od unless od is bool (then int64); a float
    takes od if od is a float (else float64).  true_divide's loops are float,
    so an integer or bool target becomes float64 (lib.rs:8125)."""
    if kind == "b":
        t = od
    elif kind == "i":
        t = _I64 if od.kind == "b" else od
    else:
        t = od if od.kind == "f" else _F64
    if code == _lib.TRUE_DIVIDE and t.kind in "biu":
        t = _F64
    return t


def _int_fits(v, t):
    info = np.iinfo(t)
    return info.min <= v <= info.max


def _weak_int_overflow(v, t):
    """numpy's OverflowError for a Python int outside dtype t: the value is
    first converted to a C long (for uint32 / uint64 targets: a C long or a
    C unsigned long), which fails beyond that range; else it is out of t's
    bounds."""
    wide_unsigned = t.kind == "u" and t.itemsize >= 4
    if not (-2 ** 63 <= v < 2 ** 63 or (wide_unsigned and 0 <= v < 2 ** 64)):
        return OverflowError("Python int too large to convert to C long")
    return OverflowError(f"Python integer {v} out of bounds for {t}")


def _weak_float(v, t):
    """A Python float as a 0-d host array of float dtype t; the conversion's
    IEEE flags (overflow, underflow) are reported under 'cast' by the
    np.errstate modes of this package."""
    with np.errstate(all="ignore"):
        r = np.array(v, dtype=t)
    if t != _F64 and _fpe_wanted():
        rv = float(r)
        fl = 0
        if math.isfinite(v) and math.isinf(rv):
            fl |= 2
        if v != 0 and math.isfinite(v) and rv != v and abs(rv) < np.finfo(t).tiny:
            fl |= 4
        if fl:
            _fpe_report(fl, "cast")
    return r


def _weak_array(x, kind, t, mode="strict"):
    """The weak scalar x materialised as a 0-d host array of dtype t.  An
    int outside an integer t: "strict" (arithmetic) raises numpy's
    OverflowError; "relaxed" (comparisons) widens to int64 / uint64 and
    compares by value (lib.rs:8146), None beyond 64 bits; "wrap" (where, in
    NumPy 2.2.6: the C cast of the concatenation family, lib.rs:8192) wraps
    to t's width and reports the cast's overflow."""
    if kind == "f" or (t.kind == "f" and kind != "b"):
        return _weak_float(float(x), t) if t.kind == "f" else np.array(x, dtype=t)
    if kind == "b" or t.kind == "b":
        return np.array(bool(x) if t.kind == "b" else int(x), dtype=t)
    v = int(x)
    if _int_fits(v, t):
        return np.array(v, dtype=t)
    if mode == "relaxed":
        t2 = _I64 if v < 0 else _U64
        return np.array(v, dtype=t2) if _int_fits(v, t2) else None
    if mode == "wrap" and -2 ** 63 <= v < 2 ** 64:
        if _fpe_wanted():
            _fpe_report(2, "cast")
        return np.array(v % (1 << (8 * t.itemsize)), dtype=np.dtype(f"u{t.itemsize}")).view(t)
    raise _weak_int_overflow(v, t)


def _default_weak(x, kind, relaxed=False):
    """Both operands weak: each takes its kind's default dtype (an int
    beyond int64 fits uint64 only in a comparison)."""
    if kind == "b":
        return np.array(x, dtype=_BOOL)
    if kind == "f":
        return np.array(x, dtype=_F64)
    v = int(x)
    if _int_fits(v, _I64):
        return np.array(v, dtype=_I64)
    if relaxed and _int_fits(v, _U64):
        return np.array(v, dtype=_U64)
    raise OverflowError("Python int too large to convert to C long")


_CMP_CONST = {  # op(array value, huge Python int v): known from the sign of v
    "less": lambda v: v > 0, "less_equal": lambda v: v > 0, "greater": lambda v: v < 0,
    "greater_equal": lambda v: v < 0, "equal": lambda v: False, "not_equal": lambda v: True}
_CMP_SWAP = {"less": "greater", "less_equal": "greater_equal", "greater": "less",
             "greater_equal": "less_equal", "equal": "equal", "not_equal": "not_equal"}


def _weak_operands(a, b, code=None, loop=None, wrap=False):
    """Host half of a binary call's operands: weak Python scalars become
    0-d host arrays (numpy's OverflowError for an int they cannot hold);
    arrays pass through.  loop: the dtype= loop, into which weak scalars
    are converted directly.  Returns (a, b, const): const is the bool
    result of a comparison against a Python int beyond 64 bits (a and b
    are then the array and None).  wrap: where()'s wrapping int cast."""
    ka, kb = _scalar_kind(a), _scalar_kind(b)
    cmp = code is not None and code >= _lib.LESS and loop is None
    if ka and kb:
        if loop is not None:
            return _weak_array(a, ka, loop), _weak_array(b, kb, loop), None
        return _default_weak(a, ka, cmp), _default_weak(b, kb, cmp), None
    # comparisons against an integer array compare a Python int by value
    relaxed = cmp and (ka or kb) and np.dtype((b if ka else a).dtype).kind in "iu"
    if ka or kb:
        arr = b if ka else a
        x, k = (a, ka) if ka else (b, kb)
        if loop is not None:
            w = _weak_array(x, k, loop)
        else:
            mode = "wrap" if wrap else "relaxed" if relaxed else "strict"
            w = _weak_array(x, k, _weak_target(k, np.dtype(arr.dtype), code), mode)
        if w is None:  # a comparison with a Python int beyond 64 bits
            name = _CODE_NAMES[code] if ka is None else _CMP_SWAP[_CODE_NAMES[code]]
            return arr, None, bool(_CMP_CONST[name](int(x)))
        return (w, arr, None) if ka else (arr, w, None)
    return a, b, None


def _operands(a, b, code=None, loop=None, wrap=False):
    """Device operands of a binary call (see _weak_operands)."""
    a, b, const = _weak_operands(a, b, code, loop, wrap)
    return asarray(a), (None if b is None else asarray(b)), const


def _resolve_stand_in(x, requested):
    """The dtype dtype= resolution checks for operand x: a weak int casts to
    any numeric loop but bool (it is converted straight into the loop's
    dtype), a weak float counts as float64, a weak bool as bool."""
    k = _scalar_kind(x)
    if k is None:
        return np.dtype(x.dtype)
    if k == "b":
        return _BOOL
    if k == "f":
        return _F64
    return requested if requested.kind != "b" else _I64


def _resolve_loop(code, a, b, dtype, casting):
    """The dtype= loop of a binary call (ionp_resolve_dtype: host only), or
    None when the call runs its promotion's loop (no dtype=, or a
    comparison, whose only loops output bool)."""
    if casting not in _lib.CASTING:
        raise ValueError("casting must be one of 'no', 'equiv', 'safe', 'same_kind', or 'unsafe'")
    if dtype is None:
        return None
    req = _dtype(dtype)
    lp = ctypes.c_int32()
    _call(_lib.load().ionp_resolve_dtype, code, _CODE[_resolve_stand_in(a, req)], _CODE[_resolve_stand_in(b, req)],
          _CODE[req], _lib.CASTING[casting], ctypes.byref(lp))
    return np.dtype(_NP_DTYPES[lp.value]) if code < _lib.LESS else None


def _finish(out):
    return out


# ---- launch-configuration policy ---------------------------------------------
_policy = None  # installed at the first call: "learned" when the model loads


def set_config_policy(policy):
    """"learned" (the default): every call runs the launch configuration the
    cost model (anionpy_amd/costmodel.json) proposes for its layout, served
    by the library's advisor (ionp_advisor_set); a layout's first call runs
    the built-in heuristic while its proposal is computed on the GPU.
    "heuristic": every call runs the built-in heuristic.  Results are
    bit-identical either way."""
    global _policy
    from . import costmodel
    if policy == "heuristic":
        costmodel.uninstall()
    elif policy == "learned":
        costmodel.install_file()
    else:
        raise ValueError(f"unknown configuration policy {policy!r}")
    _policy = policy


def get_config_policy():
    if _policy is None:
        _default_policy()
    return _policy


def _default_policy():
    try:
        set_config_policy("learned")
    except (OSError, ValueError, KeyError, IonpError):  # no usable model: the heuristic
        set_config_policy("heuristic")


def _cfg(kind, op, ins, out, axes=None, naxes=0, keepdims=False):
    """Configuration argument of a call: None, so that the library applies
    the installed policy (the advisor's proposal, else the heuristic)."""
    if _policy is None:
        _default_policy()
    return None


class Ufunc:
    """Binary ufunc object: __call__ and .reduce (ionp-py/src/lib.rs:8697-10227)."""

    nin = 2
    nout = 1

    def __init__(self, name, code, reducible):
        self.__name__ = name
        self._code = code
        self._reducible = reducible

    def __repr__(self):
        return f"<ufunc '{self.__name__}'>"

    def __call__(self, a, b, out=None, where=True, dtype=None, order="K", casting="same_kind"):
        """ufunc.__call__ (lib.rs:9204) -> binary_op (ionp-core/src/ufunc.rs:3628).

        out=: any strides; an `out` that overlaps an input other than as the
        same view gets the result through a temporary (write_out,
        ufunc.rs:10112).  where=: elements where the mask is False keep
        out's previous values.  dtype=: the loop whose output is dtype
        (ionp_resolve_dtype, lib.rs:9568-9804); the inputs are cast into it
        first (weak Python scalars straight from their value), the casts
        checked under `casting`.  order=: K (NumPy's default), C, F or A for
        a freshly allocated out.  IEEE flags of the input casts and of the
        loop are reported together under the ufunc's name."""
        wanted = _fpe_wanted()
        _fpe_arm(wanted)
        track = wanted and self._code <= _lib.TRUE_DIVIDE
        r = self._call(a, b, out, where, dtype, order, casting)
        if track:
            _fpe_check(self.__name__)
        return r

    def _call(self, a, b, out, where, dtype, order, casting):
        if isinstance(out, tuple):
            if len(out) != 1:
                raise ValueError("The 'out' tuple must have exactly one entry per ufunc output")
            out = out[0]
        L = _lib.load()
        if not _is_weak(a):
            a = asarray(a)
        if not _is_weak(b):
            b = asarray(b)
        loop = _resolve_loop(self._code, a, b, dtype, casting)
        a, b, const = _operands(a, b, self._code, loop)
        if loop is not None:  # the inputs are cast into the loop first (a buffered cast in NumPy)
            a = a if a.dtype == loop else a._astype(loop, "K")
            b = b if b.dtype == loop else b._astype(loop, "K")
        if const is not None:
            return self._constant(a, const, out, where, order)
        da, db = a._desc(), b._desc()
        t = _lib.ArrayDesc()
        _call(L.ionp_plan_binary, self._code, ctypes.byref(da), ctypes.byref(db), ctypes.byref(t))
        shape = tuple(t.shape[: t.ndim])
        masked = not (where is True)
        if out is None:
            out = empty_strided(shape, _layout(shape, t.strides[: t.ndim], order, [a, b]), _NP_DTYPES[t.dtype])
            target = out
        elif masked or _needs_temp(out, [a, b]):
            target = empty_strided(shape, list(t.strides[: t.ndim]), out.dtype)
        else:
            target = out
        do = target._desc()
        _call(L.ionp_binary, self._code, ctypes.byref(da), ctypes.byref(db), ctypes.byref(do),
              _cfg(_lib.CALL_BINARY, self._code, (da, db), do), _lib.stream_handle())
        if masked:
            self._masked_into(target, out, where)
        elif target is not out:
            _copy_into(target, out)
        return out

    def _masked_into(self, target, out, where):
        mask = asarray(where)
        if mask.dtype != np.bool_:
            raise TypeError("where= must be a boolean array")
        dm, dr, dout = mask._desc(), target._desc(), out._desc()
        _call(_lib.load().ionp_where, ctypes.byref(dm), ctypes.byref(dr), ctypes.byref(dout), ctypes.byref(dout),
              None, _lib.stream_handle())

    def _constant(self, arr, value, out, where, order):
        """A comparison whose result is known without a loop (a Python int
        beyond 64 bits against an integer array): `value` everywhere."""
        shape = arr.shape
        res = empty_strided(shape, _layout(shape, _c_strides(shape), order, [arr]), np.bool_)
        res._buf.fill_(int(value))
        if out is None:
            return res
        if not (where is True):
            self._masked_into(res, out, where)
        else:
            _copy_into(res, out)
        return out

    def accumulate(self, a, axis=0, dtype=None, out=None):
        """ufunc.accumulate -> accumulate_axis (ionp-core/src/ufunc.rs:7967):
        out[0] = a[0], out[i] = op(out[i-1], a[i]) along `axis`."""
        if not self._reducible:
            raise NotImplementedError(f"{self.__name__}.accumulate is not provided by the MI355X backend")
        return _accumulate(self._code, asarray(a), axis, dtype, out)

    def reduceat(self, a, indices, axis=0, dtype=None, out=None):
        """ufunc.reduceat (ionp-py/src/lib.rs:10127 -> reduceat_binary_axis,
        ionp-core/src/ufunc.rs:8460): per segment [indices[k], indices[k+1])
        of `axis` the reduction seeded with its first element; a
        non-increasing pair gives a[indices[k]]."""
        if not self._reducible:
            raise NotImplementedError(f"{self.__name__}.reduceat is not provided by the MI355X backend")
        return _reduceat(self._code, asarray(a), indices, axis, dtype, out)

    def reduce(self, a, axis=0, dtype=None, out=None, keepdims=False, initial=None, where=True):
        """ufunc.reduce (lib.rs:9958) -> reduce_axis (ionp-core/src/ufunc.rs:7634).
        initial=: the fold's starting value (also the result of an empty
        reduction).  where=: a mask broadcastable to `a`; only its True
        positions are reduced (maximum/minimum then need initial=)."""
        if not self._reducible:
            raise NotImplementedError(f"{self.__name__}.reduce is not provided by the MI355X backend")
        return _reduce(self._code, asarray(a), axis, dtype, keepdims, out, initial, where)


def _reduceat(code, a, indices, axis, dtype, out):
    L = _lib.load()
    idx = np.ascontiguousarray(np.asarray(indices.get() if isinstance(indices, ndarray) else indices))
    if idx.dtype.kind not in "iu" and idx.size:
        raise TypeError("Cannot cast array data from dtype('%s') to dtype('int64') according to the rule 'safe'"
                        % idx.dtype)
    idx = np.ascontiguousarray(idx.astype(np.int64).reshape(-1))
    d = a._desc()
    t = _lib.ArrayDesc()
    dt = -1 if dtype is None else _CODE[_dtype(dtype)]
    _call(L.ionp_plan_reduceat, code, ctypes.byref(d), int(axis), len(idx), dt, ctypes.byref(t))
    final = out
    if out is None or _needs_temp(out, [a]):
        out = empty_strided(tuple(t.shape[: t.ndim]), tuple(t.strides[: t.ndim]),
                            _NP_DTYPES[t.dtype] if final is None else final.dtype)
    do = out._desc()
    _call(L.ionp_reduceat, code, ctypes.byref(d), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(idx),
          int(axis), ctypes.byref(do), None, _lib.stream_handle())
    if final is not None and final is not out:
        _copy_into(out, final)
        return final
    return out


def _accumulate(code, a, axis, dtype, out):
    L = _lib.load()
    d = a._desc()
    t = _lib.ArrayDesc()
    dt = -1 if dtype is None else _CODE[_dtype(dtype)]
    _call(L.ionp_plan_accumulate, code, ctypes.byref(d), int(axis), dt, ctypes.byref(t))
    final = out
    if out is None or _needs_temp(out, [a]):
        out = empty_strided(tuple(t.shape[: t.ndim]), tuple(t.strides[: t.ndim]),
                            _NP_DTYPES[t.dtype] if final is None else final.dtype)
    do = out._desc()
    _call(L.ionp_accumulate, code, ctypes.byref(d), int(axis), ctypes.byref(do), None, _lib.stream_handle())
    if final is not None and final is not out:
        _copy_into(out, final)
        return final
    return out


def _axes_arg(axis):
    if axis is None:
        return None, -1
    if isinstance(axis, numbers.Integral):
        axis = (int(axis),)
    axis = tuple(int(a) for a in axis)
    arr = (ctypes.c_int32 * builtins.max(len(axis), 1))(*axis)
    return arr, len(axis)


def _reduce(code, a, axis, dtype, keepdims, out=None, initial=None, where=True):
    L = _lib.load()
    axes, naxes = _axes_arg(axis)
    d = a._desc()
    t = _lib.ArrayDesc()
    dt = -1 if dtype is None else _CODE[_dtype(dtype)]
    wd = None
    if not (where is True or where is None):
        wm = asarray(where)
        if wm.dtype != np.bool_:  # reduce_axis casts the mask to bool (ufunc.rs:7682)
            wm = wm.astype(np.bool_)
        wd = wm._desc()
    if wd is None:
        _call(L.ionp_plan_reduce, code, ctypes.byref(d), axes, naxes, int(bool(keepdims)), dt, ctypes.byref(t))
    else:
        _call(L.ionp_plan_reduce_where, code, ctypes.byref(d), ctypes.byref(wd), axes, naxes,
              int(bool(keepdims)), dt, ctypes.byref(t))
    final = out
    if out is None or _needs_temp(out, [a] if wd is None else [a, wm]):
        out = empty_strided(tuple(t.shape[: t.ndim]), tuple(t.strides[: t.ndim]),
                            _NP_DTYPES[t.dtype] if final is None else final.dtype)
    init = None
    if initial is not None:  # one element of the output dtype, on the host
        init = np.array(initial).astype(out.dtype).reshape(1)  # unsafe cast, as NumPy does
    do = out._desc()
    if wd is None:
        _call(L.ionp_reduce, code, ctypes.byref(d), axes, naxes, int(bool(keepdims)),
              None if init is None else init.ctypes.data, ctypes.byref(do),
              _cfg(_lib.CALL_REDUCE, code, (d,), do, axes, naxes, keepdims), _lib.stream_handle())
    else:
        _call(L.ionp_reduce_where, code, ctypes.byref(d), ctypes.byref(wd), axes, naxes, int(bool(keepdims)),
              None if init is None else init.ctypes.data, ctypes.byref(do), None, _lib.stream_handle())
    if final is not None and final is not out:
        _copy_into(out, final)
        return final
    return out


# ---- IEEE exception flags: np.seterr / np.geterr / np.errstate ----------------------
# (ionp-py/src/fpstate.rs: per-thread modes; the flags come from the device,
# see ionp_fpe_track / ionp_fpe_fetch)
_FPE_CATS = ((1, "divide", "divide by zero"), (2, "over", "overflow"), (4, "under", "underflow"),
             (8, "invalid", "invalid value"))
_FPE_MODES = ("ignore", "warn", "raise", "call", "print", "log")
_fpe_local = threading.local()


def _fpe_state():
    st = getattr(_fpe_local, "modes", None)
    if st is None:  # NumPy's defaults
        st = _fpe_local.modes = {"divide": "warn", "over": "warn", "under": "ignore", "invalid": "warn"}
        _fpe_local.call = None
    return st


def _fpe_wanted():
    return any(m != "ignore" for m in _fpe_state().values())


def _fpe_arm(on):
    """The library's flag tracking of this thread follows the modes (it is
    per thread in the library too); switching waits for the device."""
    on = bool(on)
    if getattr(_fpe_local, "armed", None) != on:
        _call(_lib.load().ionp_fpe_track, int(on))
        _fpe_local.armed = on


def _fpe_check(name):
    fl = ctypes.c_uint32()
    _call(_lib.load().ionp_fpe_fetch, _lib.stream_handle(), ctypes.byref(fl))
    if fl.value:
        _fpe_report(fl.value, name)


def _fpe_report(flags, name):
    """Warn / raise / call / print / log for each raised flag, by the modes."""
    st = _fpe_state()
    for bit, key, phrase in _FPE_CATS:
        if not flags & bit:
            continue
        mode, msg = st[key], f"{phrase} encountered in {name}"
        if mode == "warn":
            warnings.warn(msg, RuntimeWarning, stacklevel=4)
        elif mode == "raise":
            raise FloatingPointError(msg)
        elif mode == "print":
            print("Warning: " + msg, file=sys.stderr)
        elif mode == "call":
            if _fpe_local.call is None:
                raise NameError(f"call specified for {phrase} (in {name}) but no callable function found")
            _fpe_local.call(phrase, flags)
        elif mode == "log":
            if _fpe_local.call is None:
                raise NameError(f"log specified for {phrase} (in {name}) but no object with write method found.")
            _fpe_local.call.write(f"Warning: {msg}\n")


def seterr(all=None, divide=None, over=None, under=None, invalid=None):
    """numpy.seterr (ionp-py/src/fpstate.rs:143): `all` first, then the named
    categories; returns the previous modes."""
    st = _fpe_state()
    old = dict(st)
    new = dict(st)
    if all is not None:
        for k in new:
            new[k] = all
    for k, v in (("divide", divide), ("over", over), ("under", under), ("invalid", invalid)):
        if v is not None:
            new[k] = v
    for v in new.values():
        if v not in _FPE_MODES:
            raise ValueError(f"invalid error mode {v!r}")
    st.update(new)
    if getattr(_fpe_local, "armed", None) is not None:
        _fpe_arm(_fpe_wanted())  # tracking follows the new modes
    return old


def geterr():
    return dict(_fpe_state())


def seterrcall(func):
    """numpy.seterrcall: the callable ('call' mode) or object with a write
    method ('log' mode); returns the previous one."""
    _fpe_state()
    if func is not None and not callable(func) and not callable(getattr(func, "write", None)):
        raise ValueError("Only callable can be used as callback")
    old = _fpe_local.call
    _fpe_local.call = func
    return old


def geterrcall():
    _fpe_state()
    return _fpe_local.call


_UNSET = object()


class errstate(contextlib.ContextDecorator):
    """numpy.errstate (ionp-py/src/fpstate.rs:288): modes (and call=) for the
    duration of a with block."""

    def __init__(self, *, call=_UNSET, all=None, divide=None, over=None, under=None, invalid=None):
        self._call = call
        self._kw = dict(all=all, divide=divide, over=over, under=under, invalid=invalid)

    def __enter__(self):
        self._old = seterr(**self._kw)
        if self._call is not _UNSET:
            self._oldcall = seterrcall(self._call)
        return None

    def __exit__(self, *exc):
        seterr(**self._old)
        if self._call is not _UNSET:
            seterrcall(self._oldcall)
        return False


_CODE_NAMES = {_lib.LESS: "less", _lib.LESS_EQUAL: "less_equal", _lib.GREATER: "greater",
               _lib.GREATER_EQUAL: "greater_equal", _lib.EQUAL: "equal", _lib.NOT_EQUAL: "not_equal"}

add = Ufunc("add", _lib.ADD, True)
subtract = Ufunc("subtract", _lib.SUBTRACT, False)
multiply = Ufunc("multiply", _lib.MULTIPLY, True)
true_divide = Ufunc("divide", _lib.TRUE_DIVIDE, False)
divide = true_divide
maximum = Ufunc("maximum", _lib.MAXIMUM, True)
minimum = Ufunc("minimum", _lib.MINIMUM, True)
less = Ufunc("less", _lib.LESS, False)
less_equal = Ufunc("less_equal", _lib.LESS_EQUAL, False)
greater = Ufunc("greater", _lib.GREATER, False)
greater_equal = Ufunc("greater_equal", _lib.GREATER_EQUAL, False)
equal = Ufunc("equal", _lib.EQUAL, False)
not_equal = Ufunc("not_equal", _lib.NOT_EQUAL, False)


def sum(a, axis=None, dtype=None, keepdims=False, initial=None, where=True):
    """sum (ionp-py/src/reductions.rs:1032) = add.reduce with axis=None default."""
    return add.reduce(a, axis=axis, dtype=dtype, keepdims=keepdims, initial=initial, where=where)


def prod(a, axis=None, dtype=None, keepdims=False, initial=None, where=True):
    """prod (ionp-py/src/reductions.rs:1053) = multiply.reduce with axis=None
    default."""
    return multiply.reduce(a, axis=axis, dtype=dtype, keepdims=keepdims, initial=initial, where=where)


def _scan(uf, a, axis, dtype, out):
    a = asarray(a)
    if axis is None:
        a = a.reshape(-1)
        axis = 0
    return uf.accumulate(a, axis=axis, dtype=dtype, out=out)


def cumsum(a, axis=None, dtype=None, out=None):
    """cumsum (reductions.rs:1312) = add.accumulate; axis=None scans the
    C-order ravel."""
    return _scan(add, a, axis, dtype, out)


def cumprod(a, axis=None, dtype=None, out=None):
    """cumprod (reductions.rs:1324) = multiply.accumulate, axis=None as in
    cumsum."""
    return _scan(multiply, a, axis, dtype, out)


def amax(a, axis=None, keepdims=False, initial=None, where=True):
    """max / amax (reductions.rs:1114) = maximum.reduce, axis=None default."""
    return maximum.reduce(a, axis=axis, keepdims=keepdims, initial=initial, where=where)


def amin(a, axis=None, keepdims=False, initial=None, where=True):
    return minimum.reduce(a, axis=axis, keepdims=keepdims, initial=initial, where=where)


max = amax  # noqa: A001  (NumPy-named API)
min = amin  # noqa: A001


def _argext(ismax, a, axis, keepdims):
    a = asarray(a)
    L = _lib.load()
    ax = _lib.AXIS_NONE if axis is None else int(axis)
    d = a._desc()
    t = _lib.ArrayDesc()
    _call(L.ionp_plan_argext, int(ismax), ctypes.byref(d), ax, int(bool(keepdims)), ctypes.byref(t))
    out = empty_strided(tuple(t.shape[: t.ndim]), tuple(t.strides[: t.ndim]), np.int64)
    do = out._desc()
    _call(L.ionp_argext, int(ismax), ctypes.byref(d), ax, int(bool(keepdims)), ctypes.byref(do), None,
          _lib.stream_handle())
    return out


def argmax(a, axis=None, keepdims=False):
    """argmax (reductions.rs:1186) -> argext (ionp-core/src/ufunc.rs:8131)."""
    return _argext(True, a, axis, keepdims)


def argmin(a, axis=None, keepdims=False):
    return _argext(False, a, axis, keepdims)


def where(condition, x, y):
    """where (reductions.rs:3316) -> where_select (ionp-core/src/sort.rs:1869);
    a weak Python scalar takes the other operand's dtype within its kind (an
    int outside an integer dtype wraps, as NumPy 2.2.6's where casts it)."""
    c = asarray(condition)
    x, y, _ = _operands(x, y, wrap=True)
    L = _lib.load()
    dc, dx, dy = c._desc(), x._desc(), y._desc()
    t = _lib.ArrayDesc()
    _call(L.ionp_plan_where, ctypes.byref(dc), ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(t))
    out = empty_strided(tuple(t.shape[: t.ndim]), tuple(t.strides[: t.ndim]), _NP_DTYPES[t.dtype])
    do = out._desc()
    _call(L.ionp_where, ctypes.byref(dc), ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(do),
          _cfg(_lib.CALL_WHERE, 0, (dc, dx, dy), do), _lib.stream_handle())
    return out
