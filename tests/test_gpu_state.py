"""Library state across calls: IEEE flag tracking per thread (np.errstate,
ionp-py/src/fpstate.rs:103 keeps it thread-local), a batch's captured graph
against later growth of the library's scratch, and the in-launch fold of a
one-output add.reduce giving up its bounded wait."""
import threading
import warnings

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def test_errstate_leaves_no_stale_flags(ap):
    """A flag raised while tracking was on never reaches a later call: the
    modes are restored at the end of an errstate block, and batches run the
    plain kernels whatever the thread's tracking."""
    from anionpy_amd.batch import Batch, binary_call
    big = ap.asarray(np.full(3000, 1e300))
    small = ap.asarray(np.ones(3000))
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        with ap.errstate(all="warn"):
            ap.multiply(big, big)
    assert any("overflow encountered in multiply" in str(x.message) for x in w)
    with Batch([binary_call("multiply", big, big)]) as bt:
        bt.run("per_call")
        bt.run("graph")
        torch.cuda.synchronize()
    with ap.errstate(over="raise"):
        ap.add(small, small)  # clean: must not raise
        with Batch([binary_call("multiply", big, big)]) as bt:
            bt.run("per_call")  # overflows, but a batch is never tracked
            bt.run("graph")
        ap.add(small, small)
        with pytest.raises(FloatingPointError, match="overflow encountered in multiply"):
            ap.multiply(big, big)
        ap.add(small, small)
        # a comparison under tracking runs the plain kernel and leaves no flag
        ap.less(big, small)
        ap.add(small, small)


def test_flag_tracking_is_per_thread(ap):
    """Two threads with different modes on one GPU: each sees only its own
    calls' flags."""
    big = ap.asarray(np.full(4096, 1e300))
    small = ap.asarray(np.ones(4096))
    rounds = 20
    barrier = threading.Barrier(2)
    result = {}

    def overflowing():
        torch.cuda.set_device(0)
        raised = 0
        with ap.errstate(over="raise"):
            for _ in range(rounds):
                barrier.wait(timeout=60)
                try:
                    ap.multiply(big, big)
                except FloatingPointError:
                    raised += 1
        result["overflowing"] = raised

    def clean():
        torch.cuda.set_device(0)
        raised = 0
        with ap.errstate(all="raise"):
            for _ in range(rounds):
                barrier.wait(timeout=60)
                try:
                    ap.add(small, small)
                except FloatingPointError:
                    raised += 1
        result["clean"] = raised

    ts = [threading.Thread(target=overflowing), threading.Thread(target=clean)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts)
    assert result == {"overflowing": rounds, "clean": 0}


def _outputs(specs):
    torch.cuda.synchronize()
    return [s.out.get().copy() for s in specs]


def test_graph_owns_its_scratch(ap):
    """GRAPH replays after the library's scratch grew (its old buffer freed)
    and after a float32 in-launch fold refilled the library's fold slots
    with float32 sentinels: still the per-call bytes."""
    from anionpy_amd.batch import Batch, reduce_call
    from anionpy_amd.layouts import uneven_batch
    rng = np.random.default_rng(31)
    specs = uneven_batch(seed=12, n=120, max_log2=16)
    # one-output sums with in-launch folders (float64 and float32), and a
    # strided sum folded by pw_fold from library scratch
    specs.append(reduce_call(ap.asarray(rng.uniform(-1, 1, 300000)), axis=0))
    specs.append(reduce_call(ap.asarray(rng.uniform(-1, 1, (2, 90000)).astype(np.float32)), axis=1))
    specs.append(reduce_call(ap.asarray(rng.uniform(-1, 1, 500000))[::3], axis=0))
    x32 = ap.asarray(rng.uniform(-1, 1, 2 ** 23).astype(np.float32))
    x64 = ap.asarray(rng.uniform(-1, 1, 2 ** 22))
    with Batch(specs) as bt:
        bt.run("per_call")
        ref = _outputs(specs)
        for s in specs:
            s.out._buf.fill_(0xCD)
        bt.run("graph")
        assert all(a.tobytes() == b.tobytes() for a, b in zip(ref, _outputs(specs)))
        # grow the library's partials buffer (windows of 16 elements: 2^18
        # partials of 8 bytes), then a float32 one-output folded sum
        old = ap.setbufsize(16)
        try:
            ap.add.reduce(x64, axis=0).get()
        finally:
            ap.setbufsize(old)
        ap.add.reduce(x32, axis=0).get()
        for rep in range(3):
            for s in specs:
                s.out._buf.fill_(0xCD)
            bt.run("graph")
            got = _outputs(specs)
            bad = [i for i, (a, b) in enumerate(zip(ref, got)) if a.tobytes() != b.tobytes()]
            assert not bad, f"replay {rep}: calls {bad[:5]} differ from per_call"
            ap.add.reduce(x32, axis=0).get()  # library slots float32 again between replays


def test_fold_give_up_is_reported(ap):
    """A fold that gives up its bounded wait (forced from the host: window 5
    never counts as filled, 1000 polls) yields an error at the next check,
    every kernel exits, and the following sums are exact again (the slots a
    late producer wrote are refilled first)."""
    L = ap._lib.load()
    x = np.random.default_rng(32).uniform(-1, 1, 2 ** 20)
    X = ap.asarray(x)
    want = np.asarray(O.reduce("add", x, None))
    assert ap.add.reduce(X, axis=0).get().tobytes() == want.tobytes()
    assert L.ionp_debug_fold_stall(5, 1000) == 0
    try:
        r = ap.add.reduce(X, axis=0)
        with pytest.raises(RuntimeError, match="gave up waiting"):
            r.get()
    finally:
        L.ionp_debug_fold_stall(-1, 1 << 26)
    for _ in range(3):
        assert ap.add.reduce(X, axis=0).get().tobytes() == want.tobytes()
