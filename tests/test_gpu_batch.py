"""Batched execution (persistent kernel, hipGraph, per-call) of an uneven
batch: every mode returns the same bytes, equal to the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def expected(spec):
    ins = [x.get() for x in spec.inputs]
    from anionpy_amd import _lib
    if spec.kind == _lib.CALL_BINARY:
        name = {0: "add", 2: "multiply", 6: "less"}[spec.op]
        return O.binary(name, ins[0], ins[1])
    if spec.kind == _lib.CALL_WHERE:
        return O.where(*ins)
    axes = None if spec.rec.naxes < 0 else tuple(spec.rec.axes[i] for i in range(spec.rec.naxes))
    return O.reduce("add", ins[0], axes)


def test_uneven_batch_modes_agree_with_oracle(ap):
    from anionpy_amd.batch import Batch
    from anionpy_amd.layouts import uneven_batch
    specs = uneven_batch(seed=11, n=300, max_log2=14)
    outs = {}
    with Batch(specs) as bt:
        info = bt.info()
        assert info["calls"] == 300 and info["tiles"] > 0
        for mode in ("per_call", "persistent", "graph", "persistent", "graph"):
            for s in specs:
                s.out._buf.fill_(0xAB)  # poison: every output byte must be rewritten
            bt.run(mode)
            torch.cuda.synchronize()
            got = [s.out.get().copy() for s in specs]
            if mode in outs:
                assert all(a.tobytes() == b.tobytes() for a, b in zip(outs[mode], got)), f"{mode} not repeatable"
            outs[mode] = got
    for mode in ("persistent", "graph"):
        for i, (a, b) in enumerate(zip(outs["per_call"], outs[mode])):
            assert a.tobytes() == b.tobytes(), f"{mode} differs from per_call at call {i}"
    bad = []
    for i, s in enumerate(specs):
        want = np.asarray(expected(s))
        got = outs["persistent"][i]
        if got.shape != want.shape or np.ascontiguousarray(got).tobytes() != np.ascontiguousarray(want).tobytes():
            bad.append(i)
    assert not bad, f"{len(bad)} calls differ from the oracle, first {bad[:5]}"


def test_uneven_batch_full_size_matches_oracle(ap):
    """The bench's batch at its stated size (10,000 calls from seed 5, sizes
    log-uniform in 16..2^20 elements): PERSISTENT and GRAPH outputs, every
    byte, against the C oracle on host copies -- including the long narrow
    chains and multi-window wide sums inside the persistent kernel."""
    from anionpy_amd.batch import Batch
    from anionpy_amd.layouts import uneven_batch
    specs = uneven_batch(seed=5, n=10000)
    with Batch(specs) as bt:
        for mode in ("persistent", "graph"):
            for s in specs:
                s.out._buf.fill_(0xAB)
            bt.run(mode)
            torch.cuda.synchronize()
            bad = []
            for i, s in enumerate(specs):
                want = np.asarray(expected(s))  # 0-d outputs stay 0-d (ascontiguousarray would make them 1-d)
                got = s.out.get()
                if got.shape != want.shape or np.ascontiguousarray(got).tobytes() != np.ascontiguousarray(want).tobytes():
                    bad.append(i)
            print(f"{mode}: {len(specs) - len(bad)}/{len(specs)} calls equal to the oracle")
            assert not bad, f"{mode}: {len(bad)} calls differ from the oracle, first {bad[:5]}"


def test_float16_calls_in_a_batch(ap):
    """float16 calls mixed into a batch (single-call launches beside the
    persistent kernel): every mode, every byte against the oracle."""
    from anionpy_amd.batch import Batch, binary_call, reduce_call, where_call
    rng = np.random.default_rng(21)
    h = lambda *s: rng.standard_normal(s).astype(np.float16) * 10  # noqa: E731
    xs = [h(300, 70), h(70), h(5000), h(40, 9000)]
    specs = [binary_call("add", xs[0], xs[1]), binary_call("multiply", xs[0], rng.standard_normal((300, 1))),
             binary_call("less", xs[1], xs[1][::-1]), reduce_call(xs[2], axis=0), reduce_call(xs[3], axis=1),
             reduce_call(xs[3], axis=0, dtype=np.float32),
             where_call(rng.random((300, 70)) < 0.5, xs[0], np.float16(2.5)),
             binary_call("add", rng.standard_normal((64, 64)), rng.standard_normal(64))]
    with Batch(specs) as bt:
        ref = None
        for mode in ("per_call", "persistent", "graph", "persistent"):
            for s in specs:
                s.out._buf.fill_(0xAB)
            bt.run(mode)
            torch.cuda.synchronize()
            got = [s.out.get().copy() for s in specs]
            if ref is None:
                ref = got
            assert all(a.tobytes() == b.tobytes() for a, b in zip(ref, got)), mode
    for i, s in enumerate(specs):
        ins = [x.get() for x in s.inputs]
        if s.kind == ap._lib.CALL_BINARY:
            want = O.binary({0: "add", 2: "multiply", 6: "less"}[s.op], ins[0], ins[1])
        elif s.kind == ap._lib.CALL_WHERE:
            want = O.where(*ins)
        else:
            x = ins[0] if s.out.dtype == np.float16 else O.cast(ins[0], s.out.dtype)
            want = O.reduce("add", x, s.rec.axes[0], dtype=s.out.dtype)
        assert np.asarray(ref[i]).tobytes() == np.ascontiguousarray(want).tobytes(), i


def test_batch_rejects_bad_calls(ap):
    from anionpy_amd.batch import Batch, reduce_call
    x = ap.asarray(np.zeros((4, 5)))
    with pytest.raises(NotImplementedError):
        Batch([reduce_call(x, axis=1, op="maximum")])


def test_batch_autotuner_schedules_are_bit_identical(ap):
    """ionp_batch_autotune times the persistent run under each tile schedule;
    the tuned batch (and every schedule it tried) writes the same bytes as
    the default schedule."""
    from anionpy_amd.batch import BATCH_GRID, Batch, autotune
    from anionpy_amd.layouts import uneven_batch
    specs = uneven_batch(seed=12, n=400, max_log2=16)
    with Batch(specs) as bt:
        bt.run("persistent")
        torch.cuda.synchronize()
        want = [s.out.get().copy() for s in specs]
    grid = [BATCH_GRID[0], (8192, 4096, 1), (65536, 16384, 8), (32768, 8192, 2)]
    t = autotune(specs, grid=grid, reps=2)
    assert t.shape == (4,) and (t > 0).all()
    # the last schedule timed left its outputs in place
    assert all(a.tobytes() == s.out.get().tobytes() for a, s in zip(want, specs))
    bt = Batch.tuned(specs, grid=grid, reps=2)
    try:
        assert bt.schedule in grid and bt.tune_times.shape == (4,)
        for s in specs:
            s.out._buf.fill_(0xAB)
        bt.run("persistent")
        torch.cuda.synchronize()
        assert all(a.tobytes() == s.out.get().tobytes() for a, s in zip(want, specs))
    finally:
        bt.close()
    # an ew_tile that is not a multiple of 1024 is a schedule the batch
    # cannot use: reported as a negative time
    assert autotune(specs, grid=[(1000, 8192, 4)], reps=1)[0] < 0
