"""CPU oracle of the warm-start network -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` import this module, as the checker of the HIP network kernels
(``grg-spmv_amd/csrc/grg_mlp.hip``). The network has no counterpart in the
reference; the oracle is the textbook dense network in FP64 numpy:
y = W x + b with ReLU between layers, weighted squared-error loss, Adam
(Kingma & Ba 2015, Algorithm 1). It is pinned by central finite differences
(``tests/test_lmm_cpu.py``).
"""

from __future__ import annotations

import numpy as np


def forward(Ws, bs, X):
    """Activations [X, h1, ..., out]; ReLU after every layer but the last."""
    acts = [np.asarray(X, dtype=np.float64)]
    for li, (W, b) in enumerate(zip(Ws, bs)):
        z = acts[-1] @ np.asarray(W, dtype=np.float64).T + np.asarray(b, dtype=np.float64)
        if li < len(Ws) - 1:
            z = np.maximum(z, 0.0)
        acts.append(z)
    return acts


def loss_and_grads(Ws, bs, X, T, w=None):
    """Loss (1/B) sum_b sum_j w_j (y - t)^2 and its gradients for every W and b."""
    acts = forward(Ws, bs, X)
    B = acts[0].shape[0]
    wj = np.ones(acts[-1].shape[1]) if w is None else np.asarray(w, dtype=np.float64)
    d = acts[-1] - np.asarray(T, dtype=np.float64)
    loss = float((wj * d * d).sum() / B)
    g = 2.0 * wj * d / B
    dWs, dbs = [None] * len(Ws), [None] * len(Ws)
    for li in range(len(Ws) - 1, -1, -1):
        if li < len(Ws) - 1:
            g = g * (acts[li + 1] > 0.0)
        dWs[li] = g.T @ acts[li]
        dbs[li] = g.sum(axis=0)
        g = g @ np.asarray(Ws[li], dtype=np.float64)
    return loss, dWs, dbs


def adam(p, g, m, v, t, *, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    p = p - lr * (m / (1.0 - beta1 ** t)) / (np.sqrt(v / (1.0 - beta2 ** t)) + eps)
    return p, m, v


__all__ = ["adam", "forward", "loss_and_grads"]
