"""Synthetic phenotypes y = X_s beta + e over the GRGs of a standardized operator.

For each phenotype: h2 ~ U[h2_lo, h2_hi]; a fraction ``causal_frac`` of the
model variants is causal with beta_j ~ N(0, h2 / m_causal) on the standardized
genotypes; e ~ N(0, 1 - h2); delta = (1 - h2) / h2 * xfro2 / (m_proj * n), the
log-delta / h2 relation of the reference (``scripts/bolt_lmm_inf/core.py:1356-1357``,
``log_delta_from_h2``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Phenotypes:
    y: object        # n x c FP64 device tensor
    h2: np.ndarray   # c
    delta: object    # c FP64 device tensor


def simulate(op, count: int, *, seed: int, causal_frac: float = 0.01, h2_range=(0.1, 0.6)) -> Phenotypes:
    """``count`` phenotypes, generated ``op.k`` at a time."""
    import torch

    rng = np.random.default_rng(seed)
    sizes = [s.shape[0] for s in op.scales]
    model = [np.flatnonzero(s > 0.0) for s in op.scales]
    offsets = np.cumsum([0] + [m.size for m in model])
    m_causal = max(1, int(round(causal_frac * op.m_proj)))
    ys, h2s = [], []
    for start in range(0, int(count), op.k):
        c = min(op.k, int(count) - start)
        h2 = rng.uniform(h2_range[0], h2_range[1], size=c)
        betas = [np.zeros((sz, c)) for sz in sizes]
        for col in range(c):
            picks = rng.choice(op.m_proj, size=m_causal, replace=False)
            effects = rng.standard_normal(m_causal) * np.sqrt(h2[col] / m_causal)
            g_of = np.searchsorted(offsets, picks, side="right") - 1
            for g in np.unique(g_of):
                sel = g_of == g
                betas[g][model[g][picks[sel] - offsets[g]], col] = effects[sel]
        genetic = op.genetic_values(betas)
        noise = torch.from_numpy(rng.standard_normal((op.n, c)) * np.sqrt(1.0 - h2)[None, :]).to(op.device)
        ys.append(genetic + noise)
        h2s.append(h2)
    h2 = np.concatenate(h2s)
    delta = (1.0 - h2) / h2 * op.xfro2 / (op.m_proj * op.n)
    return Phenotypes(y=torch.cat(ys, dim=1), h2=h2, delta=torch.from_numpy(delta).to(op.device))


__all__ = ["Phenotypes", "simulate"]
