"""Every golden vector through the HIP backend (bit-exact vs NumPy 2.2.6)."""
import numpy as np
import pytest
import torch

from fixture_cases import args_axis, check_result, host_input, load, py_scalar

pytestmark = pytest.mark.gpu

KINDS = ["binary", "where", "cast", "reduce", "argext", "accumulate", "reduceat", "fpe"]

EXC = {"ValueError": ValueError, "TypeError": TypeError, "IndexError": IndexError,
       "AxisError": np.exceptions.AxisError, "UFuncTypeError": TypeError, "OverflowError": OverflowError}


def device_input(ap, spec, npz, cache):
    """Rebuild the exact operand view on the device (same offset/strides);
    a Python scalar operand stays the scalar."""
    if "py" in spec:
        return py_scalar(spec)
    if "gen" in spec:
        return ap.asarray(host_input(spec, npz))
    buf = npz[spec["buf"]]
    key = spec["buf"]
    if key not in cache:
        cache[key] = torch.from_numpy(np.ascontiguousarray(buf).view(np.uint8).copy()).cuda()
    return ap.ndarray(cache[key], buf.dtype, spec["shape"], spec["strides"], spec["offset"])


class _Host:
    """A host result where the other kinds return device arrays."""

    def __init__(self, a):
        self.a = a

    def get(self):
        return self.a


def run_device(ap, case, npz, cache):
    ins = [device_input(ap, s, npz, cache) for s in case["inputs"]]
    k, op, args = case["kind"], case["op"], case["args"]
    if k == "binary":
        kw = {}
        if "dtype" in args:
            kw["dtype"] = np.dtype(args["dtype"])
        if "casting" in args:
            kw["casting"] = args["casting"]
        r = getattr(ap, op)(ins[0], ins[1], **kw)
        return r
    if k == "where":
        return ap.where(*ins)
    if k == "cast":
        return ins[0].astype(np.dtype(args["dtype"]))
    if k == "reduce":
        return getattr(ap, op).reduce(ins[0], axis=args_axis(args), dtype=args.get("dtype"),
                                      keepdims=bool(args.get("keepdims", False)), initial=args.get("initial"),
                                      where=ins[1] if args.get("where") else True)
    if k == "fpe":
        got = []
        with ap.errstate(all="call", call=lambda t, f: got.append(f)):
            if op == "cast":
                ins[0].astype(np.dtype(args["dtype"]))
            else:
                getattr(ap, op)(ins[0], ins[1])
        r = 0
        for f in got:
            r |= f
        return _Host(np.array([r], dtype=np.int64))
    if k == "reduceat":
        return getattr(ap, op).reduceat(ins[0], args["indices"], axis=args["axis"], dtype=args.get("dtype"))
    if k == "accumulate":
        if op in ("cumsum", "cumprod"):
            return getattr(ap, op)(ins[0], axis=args.get("axis"))
        return getattr(ap, op).accumulate(ins[0], axis=args["axis"], dtype=args.get("dtype"))
    return getattr(ap, op)(ins[0], axis=args.get("axis"), keepdims=bool(args.get("keepdims", False)))


# every kind with IEEE flag tracking off (the plain kernels), and the kinds
# whose kernels have flag variants once more with tracking on (same bytes)
RUNS = [(k, False) for k in KINDS] + [("binary", True), ("cast", True)]


@pytest.mark.parametrize("kind,tracked", RUNS, ids=[k + ("-tracked" if t else "") for k, t in RUNS])
def test_golden_on_gpu(ap, kind, tracked):
    cases, npz = load()
    cache = {}
    bad = []
    n = 0
    ctx = ap.errstate(all="call", call=lambda t, f: None) if tracked else ap.errstate(all="ignore")
    for c in cases:
        if c["kind"] != kind:
            continue
        n += 1
        exp = c["expected"]
        try:
            with ctx:
                got = run_device(ap, c, npz, cache)
        except Exception as e:  # noqa: BLE001
            if "error" in exp and isinstance(e, EXC.get(exp["error"], Exception)):
                if str(e) != exp["msg"]:
                    bad.append((c["name"], f"message {str(e)!r} != NumPy's {exp['msg']!r}"))
                continue
            bad.append((c["name"], f"raised {type(e).__name__}: {e}"))
            continue
        if "error" in exp:
            bad.append((c["name"], "did not raise " + exp["error"]))
            continue
        why = check_result(c, got.get(), npz)
        if why:
            bad.append((c["name"], why))
    torch.cuda.synchronize()
    print(f"{kind}: {n - len(bad)}/{n} bit-exact")
    assert n > 0
    assert not bad, f"{len(bad)} of {n} mismatches, first: {bad[:6]}"
