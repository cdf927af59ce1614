"""The cost model's HIP kernels against the float64 CPU oracle, and what
the trained model gains over the built-in heuristic on held-out layouts."""
import numpy as np
import pytest
import torch

import oracle.mlp_ref as R

pytestmark = pytest.mark.gpu


def test_mlp_gradients_match_float64_oracle(ap):
    from anionpy_amd import costmodel as CM
    m = CM.MLP(seed=3)
    rng = np.random.default_rng(0)
    n = 37
    x = rng.uniform(0, 1, (n, 24)).astype(np.float32)
    labels = np.stack([rng.integers(0, k, n) for k in CM.HEAD_SIZES], 1).astype(np.int32)
    xt = torch.as_tensor(x, device="cuda")
    yt = torch.as_tensor(labels, device="cuda")
    loss, gW, gb = m.loss_and_grads(xt, yt)
    Ws = [w.cpu().numpy() for w in m.W]
    bs = [b.cpu().numpy() for b in m.b]
    ref_loss, rW, rb = R.loss_and_grads(x, labels, Ws, bs, CM.HEAD_OFF)
    assert abs(float(loss[0]) - ref_loss) <= 1e-5 * max(1.0, abs(ref_loss))
    for g, r in zip(gW + gb, rW + rb):
        g = g.cpu().numpy().astype(np.float64)
        scale = max(np.abs(r).max(), 1e-6)
        assert np.abs(g - r).max() <= 2e-5 * scale + 1e-7
    # forward of the trained-model path (predict) agrees with the oracle's argmax
    acts = R.forward(x, Ws, bs)
    z = acts[-1]
    want = np.stack([z[:, CM.HEAD_OFF[h]:CM.HEAD_OFF[h + 1]].argmax(1) for h in range(len(CM.HEAD_SIZES))], 1)
    got_idx = m.predict(x)
    from anionpy_amd.autotune import config_index
    assert list(got_idx) == [config_index(list(w)) for w in want]


def test_adam_step_matches_reference(ap):
    from anionpy_amd import _lib
    import ctypes
    p = torch.linspace(-1, 1, 100, device="cuda")
    g = torch.sin(torch.arange(100, device="cuda", dtype=torch.float32))
    m = torch.zeros_like(p)
    v = torch.zeros_like(p)
    p0, g0 = p.cpu().numpy().astype(np.float64), g.cpu().numpy().astype(np.float64)
    L = _lib.load()
    assert L.ionp_mlp_adam(ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(m.data_ptr()),
                           ctypes.c_void_p(v.data_ptr()), 100, 0.01, 0.9, 0.999, 1e-8, 1, _lib.stream_handle()) == 0
    mm = 0.1 * g0
    vv = 0.001 * g0 * g0
    want = p0 - 0.01 * (mm / 0.1) / (np.sqrt(vv / 0.001) + 1e-8)
    assert np.abs(p.cpu().numpy() - want).max() < 1e-5


def test_learned_config_beats_heuristic_on_held_out(ap):
    """Train on a small sweep, evaluate on unseen layouts: the model's
    choices must be faster in aggregate than the built-in heuristic."""
    from anionpy_amd import autotune, costmodel as CM
    from anionpy_amd.layouts import Pool, tuning_layouts
    pool = Pool(1 << 29)
    train = tuning_layouts(101, 800, pool, min_log2=16, max_log2=22)
    test = tuning_layouts(202, 250, pool, min_log2=16, max_log2=22)
    t_train = autotune.sweep(train, reps=2)
    t_test = autotune.sweep(test, reps=2)
    f_train = np.stack([s.features() for s in train])
    f_test = np.stack([s.features() for s in test])
    m = CM.MLP(seed=0)
    hist = m.train(f_train, CM.labels_of(t_train), epochs=60, batch=128)
    assert hist[-1] < hist[0]
    chosen = m.predict(f_test)
    learned, _ = CM.regret(t_test, chosen)
    heuristic, _ = CM.regret(t_test, np.zeros(len(test), dtype=int))
    print(f"held-out aggregate regret: learned {learned:.3f}, heuristic {heuristic:.3f}")
    assert learned < heuristic
