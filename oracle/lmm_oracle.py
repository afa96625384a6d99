"""CPU oracle of the mixed-model CG consumer -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` import this module. It restates, densely in FP64 numpy on tiny
GRGs, the reference's BOLT-LMM-inf operator and solver
(``scripts/bolt_lmm_inf/core.py``):

* ``variant_stats``: ``attach_bed_stats`` ``:935-967`` on the dense diploid
  dosages of the GRG (mean, mean-centred norm^2, scale);
* ``dense_kinship``: K = P X_s X_s^T P / M with the all-ones covariate basis
  (``apply_k`` ``:1282-1301``, ``project`` ``:1228-1234``);
* ``conj_grad``: ``bolt_conj_grad_solve`` ``:1304-1353``.

Pinned in ``tests/test_lmm_cpu.py``: K is symmetric positive semi-definite
with trace n - 1 and K 1 = 0; ``conj_grad`` agrees with a direct solve.
"""

from __future__ import annotations

import numpy as np

from oracle import grg_oracle as O


def dosages(g) -> np.ndarray:
    """Individuals x mutations dosage matrix: the oracle's up sweep of the identity by
    individual (row i = the dosages of individual i)."""
    return O.matmul(g, np.eye(g.num_individuals), "up", by_individual=True)


def variant_stats(G: np.ndarray):
    n = G.shape[0]
    sum_g = G.sum(axis=0)
    sumsq = (G * G).sum(axis=0)
    mean = sum_g / n
    norm2 = sumsq - sum_g * sum_g / n
    scale = np.where(norm2 > 0.0, np.sqrt((n - 1.0) / np.where(norm2 > 0.0, norm2, 1.0)), 0.0)
    return mean, norm2, scale


def project(V: np.ndarray) -> np.ndarray:
    return V - V.mean(axis=0, keepdims=True)


def dense_kinship(Gs) -> tuple[np.ndarray, int]:
    """K over several GRGs' dosage matrices (same individuals), and M (model variants)."""
    n = Gs[0].shape[0]
    K = np.zeros((n, n))
    m = 0
    for G in Gs:
        mean, _, scale = variant_stats(G)
        keep = scale > 0.0
        Xs = (G[:, keep] - mean[keep]) * scale[keep]
        Xs = project(Xs)
        K += Xs @ Xs.T
        m += int(keep.sum())
    return K / m, m


def conj_grad(H: np.ndarray, rhs: np.ndarray, *, rel_tol: float, max_iter: int):
    """Full-batch CG with per-iteration projection; returns (x, iterations)."""
    b = project(np.asarray(rhs, dtype=np.float64))
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    r2_orig = (r * r).sum(axis=0)
    r2_old = r2_orig.copy()
    it = 0
    for it in range(1, int(max_iter) + 1):
        hp = H @ p
        alpha = r2_old / (p * hp).sum(axis=0)
        x += p * alpha
        r -= hp * alpha
        r = project(r)
        r2_new = (r * r).sum(axis=0)
        if not np.any(np.sqrt(r2_new / r2_orig) > rel_tol):
            break
        p = p * (r2_new / r2_old) + r
        r2_old = r2_new
    return project(x), it


__all__ = ["conj_grad", "dense_kinship", "dosages", "project", "variant_stats"]
