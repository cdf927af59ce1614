"""float64 NumPy restatement of the cost-model MLP (test infrastructure
only): forward, summed-head softmax cross-entropy (hard labels or soft targets)
and backward, used to check the HIP kernels' gradients."""
import numpy as np


def forward(x, Ws, bs):
    acts = [np.asarray(x, dtype=np.float64)]
    for i, (w, b) in enumerate(zip(Ws, bs)):
        z = acts[-1] @ np.asarray(w, np.float64) + np.asarray(b, np.float64)
        acts.append(np.maximum(z, 0.0) if i < len(Ws) - 1 else z)
    return acts


def loss_and_grads(x, labels, Ws, bs, head_off):
    """labels: int class per head (n x heads), or float soft targets
    (n x total, non-negative weights over each head's values; a row whose
    head sums to m counts m times)."""
    soft = np.asarray(labels).dtype.kind == "f"
    acts = forward(x, Ws, bs)
    z = acts[-1]
    n = z.shape[0]
    dz = np.zeros_like(z)
    loss = 0.0
    for h in range(len(head_off) - 1):
        a, b = head_off[h], head_off[h + 1]
        zz = z[:, a:b] - z[:, a:b].max(axis=1, keepdims=True)
        p = np.exp(zz) / np.exp(zz).sum(axis=1, keepdims=True)
        if soft:  # labels is q[n, total]: non-negative weights per head (mass m per row)
            q = np.asarray(labels[:, a:b], np.float64)
            loss += -(q * np.log(p)).sum()
            dz[:, a:b] = (p * q.sum(axis=1, keepdims=True) - q) / n
            continue
        lab = labels[:, h]
        loss += -np.log(p[np.arange(n), lab]).sum()
        g = p.copy()
        g[np.arange(n), lab] -= 1.0
        dz[:, a:b] = g / n
    loss /= n
    gW, gb = [None] * len(Ws), [None] * len(Ws)
    dy = dz
    for i in range(len(Ws) - 1, -1, -1):
        if i < len(Ws) - 1:
            dy = dy * (acts[i + 1] > 0)
        gW[i] = acts[i].T @ dy
        gb[i] = dy.sum(axis=0)
        dy = dy @ np.asarray(Ws[i], np.float64).T
    return loss, gW, gb
