"""CPU oracle of the mixed-model CG consumer -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` import this module. It restates, densely in FP64 numpy on tiny
GRGs, the reference's BOLT-LMM-inf operator and solver
(``scripts/bolt_lmm_inf/core.py``):

* ``variant_stats``: ``attach_bed_stats`` ``:935-967`` on the dense diploid
  dosages of the GRG (mean, mean-centred norm^2, scale);
* ``dense_kinship``: K = P X_s X_s^T P / M with the all-ones covariate basis
  (``apply_k`` ``:1282-1301``, ``project`` ``:1228-1234``);
* ``conj_grad``: ``bolt_conj_grad_solve`` ``:1304-1353``;
* ``mc_scaling``: ``_compute_mc_scaling`` ``:1421-1502`` with direct solves of
  (K + delta I) z = P rhs in place of CG;
* ``loco_solutions``: ``solve_loco_hinv_y`` ``:1565-1591`` (K_-c over the other
  GRGs, divided by their model-variant count), direct solves.

Pinned in ``tests/test_lmm_cpu.py``: K is symmetric positive semi-definite
with trace n - 1 and K 1 = 0; ``conj_grad`` agrees with a direct solve; the
score square sums equal M z^T K z; f_REML is 0 at every delta when K is a
multiple of P; a data column equal to trial t gives f_REML = the trial's
as-data value; a LOCO solve over two GRGs equals the full solve on the other
GRG's own kinship.
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


def standardized(Gs):
    """Per GRG: (P X_s on the model variants (n x M_g), the model-variant mask)."""
    out = []
    for G in Gs:
        mean, _, scale = variant_stats(G)
        keep = scale > 0.0
        out.append((project((G[:, keep] - mean[keep]) * scale[keep]), keep))
    return out


def kinship_of(Xs_list) -> np.ndarray:
    n = Xs_list[0].shape[0]
    m = sum(X.shape[1] for X in Xs_list)
    K = np.zeros((n, n))
    for X in Xs_list:
        K += X @ X.T
    return K / max(m, 1)


def score_square_sums(Xs_list, Z: np.ndarray) -> np.ndarray:
    """sum over GRGs and model variants of (X_s^T P z)^2, per column of Z."""
    PZ = project(Z)
    return sum(((X.T @ PZ) ** 2).sum(axis=0) for X in Xs_list)


def genetic_values(Xs_list, masks, betas) -> np.ndarray:
    """P X_s w for per-GRG effects (M_g x c, rows of non-model variants ignored)."""
    return project(sum(X @ b[keep] for X, keep, b in zip(Xs_list, masks, betas)))


def mc_scaling(Xs_list, G: np.ndarray, E: np.ndarray, y: np.ndarray, log_delta: float) -> dict:
    n = Xs_list[0].shape[0]
    K = kinship_of(Xs_list)
    delta = float(np.exp(log_delta))
    rhs = np.concatenate([project(E * np.sqrt(delta) + G), project(y.reshape(n, 1))], axis=1)
    Z = project(np.linalg.solve(K + delta * np.eye(n), rhs))
    beta = score_square_sums(Xs_list, Z)
    eps = (Z * Z).sum(axis=0)
    t = G.shape[1]
    rb, re = beta[:t].sum(), eps[:t].sum()
    ratio = beta[t] / eps[t]
    f = float(np.log(ratio / (rb / re)))
    jacks = [float(np.log(ratio / ((rb - beta[i]) / (re - eps[i])))) for i in range(t)] + [f]
    as_data = [float(np.log((beta[i] / eps[i]) / (rb / re))) for i in range(t)]
    sigma2_k = float(project(y.reshape(n, 1))[:, 0] @ Z[:, t]) / (n - 1)
    return {"f_reml": f, "f_jacks": jacks, "f_rands_as_data": as_data, "sigma2_k": sigma2_k, "hinv_y": Z[:, t],
            "beta": beta, "eps": eps}


def loco_solutions(Xs_list, y: np.ndarray, delta: float) -> np.ndarray:
    """n x C: column c solves (K_-c + delta I) z = P y."""
    n = Xs_list[0].shape[0]
    Py = project(y.reshape(n, 1))[:, 0]
    cols = []
    for c in range(len(Xs_list)):
        rest = [X for g, X in enumerate(Xs_list) if g != c]
        K = kinship_of(rest) if rest and sum(X.shape[1] for X in rest) > 0 else np.zeros((n, n))
        cols.append(np.linalg.solve(K + float(delta) * np.eye(n), Py))
    return project(np.stack(cols, axis=1))


__all__ = ["conj_grad", "dense_kinship", "dosages", "genetic_values", "kinship_of", "loco_solutions", "mc_scaling", "project",
           "score_square_sums", "standardized", "variant_stats"]
