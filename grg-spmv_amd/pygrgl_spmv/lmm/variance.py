"""Variance components and leave-one-chromosome-out solves (BOLT-LMM-inf numerics).

Restates the reference's BOLT-LMM-inf script on the HIP sweeps
(``scripts/bolt_lmm_inf/core.py``):

* ``mc_components``: ``_generate_bolt_mc_components`` ``:1373-1418``: per trial,
  random effects g = P X_s w with w ~ N(0, 1/M) on every model variant, noise
  e = P N(0, I), and P y;
* ``mc_scaling``: ``_compute_mc_scaling`` ``:1421-1502``: one batched CG solve of
  (K + delta I) z = P (e sqrt(delta) + g) for every trial and P y, then
  f_REML(log delta) = log((||X_s^T z_y||^2 / ||z_y||^2) / (sum_t ||X_s^T z_t||^2 /
  sum_t ||z_t||^2)), its jackknife values, and sigma_K^2 = y^T z_y / dim;
* ``fit_variance_components``: ``fit_bolt_variance_components`` ``:1505-1562``:
  the secant search on log delta, starting from h2 = 0.25 and 0.125 or 0.5;
* ``solve_loco``: ``solve_loco_hinv_y`` ``:1565-1591``: for every GRG c, (K_-c +
  delta I) z = P y with K_-c the kinship over the other GRGs, all solved in one
  batched CG whose columns each use their own operator.

The covariate basis is the all-ones vector (P = centring), so dim = n - 1 and
xfro2 = M (n - 1). The random draws come from numpy's generator seeded by
``seed``; the reference reproduces BOLT's own Mersenne-Twister and ziggurat
streams to match its binary draw for draw, which this build does not.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from pygrgl_spmv.lmm.cg import conj_grad
from pygrgl_spmv.lmm.operator import center


@dataclass
class McScaling:
    log_delta: float
    f_reml: float
    f_jacks: tuple
    f_rands_as_data: tuple
    sigma2_k: float
    hinv_y: object        # n FP64 device tensor, (K + delta I)^-1 P y
    iterations: int


@dataclass
class VarianceFit:
    log_delta: float
    delta: float
    sigma_g2: float
    sigma_e2: float
    h2: float
    hinv_y: object
    evaluations: tuple    # the McScaling of every log delta tried, in order


def _dim(op) -> int:
    return int(op.n) - 1


def log_delta_from_h2(op, h2: float) -> float:
    """``core.py:1356-1357``."""
    return math.log(float(op.xfro2) / (float(op.m_proj) * _dim(op)) * (1.0 - float(h2)) / float(h2))


def h2_from_log_delta(op, log_delta: float) -> float:
    """``core.py:1360-1361``."""
    return float(op.xfro2) / (float(op.xfro2) + float(op.m_proj) * _dim(op) * math.exp(float(log_delta)))


def score_square_sums(op, Z):
    """sum over GRGs and model variants of (X_s^T P z)^2 per column of Z (n x c), as a
    c FP64 device tensor (``_sum_score_squares`` ``core.py:1364-1370``)."""
    import torch

    c = int(Z.shape[1])
    total = torch.zeros((c,), dtype=torch.float64, device=op.device)
    for sc in op.scores(Z):
        total += (sc * sc).sum(dim=0)
    return total


def mc_draws(scales, n: int, m_proj: int, *, trials: int, seed: int):
    """The random inputs of the MC scaling: per GRG an M_g x trials matrix of effects
    w ~ N(0, 1 / M) on the model variants (0 elsewhere), and an n x trials matrix of
    N(0, 1) noise, both FP64 numpy from numpy's generator seeded by ``seed + 1``."""
    rng = np.random.default_rng(int(seed) + 1)
    inv_sqrt_m = 1.0 / math.sqrt(float(m_proj))
    betas = []
    for scale in scales:
        w = rng.standard_normal((np.asarray(scale).size, int(trials))) * inv_sqrt_m
        w[np.asarray(scale) <= 0.0] = 0.0
        betas.append(w)
    noise = rng.standard_normal((int(n), int(trials)))
    return betas, noise


def mc_components(op, y, *, trials: int, seed: int):
    """(G, E, P y) on the device (``_generate_bolt_mc_components`` ``core.py:1373-1418``):
    G = P X_s w per trial, E = P noise, both n x trials FP64."""
    import torch

    betas, noise = mc_draws(op.scales, op.n, op.m_proj, trials=trials, seed=seed)
    G = torch.zeros((op.n, int(trials)), dtype=torch.float64, device=op.device)
    for start in range(0, int(trials), op.k):
        stop = min(int(trials), start + op.k)
        G[:, start:stop] = op.genetic_values([b[:, start:stop] for b in betas])
    E = center(torch.from_numpy(noise).to(op.device))
    ydev = center(torch.as_tensor(np.asarray(y, dtype=np.float64).reshape(op.n, -1)[:, :1]).to(op.device))
    return G, E, ydev


def reml_objective(data_beta: float, data_eps: float, rand_beta, rand_eps):
    """(f_REML, jackknife values (one per left-out trial, then f_REML), per-trial
    values as if each trial were the data), from the score and residual square sums
    (``core.py:1471-1492``)."""
    rand_beta = [float(v) for v in rand_beta]
    rand_eps = [float(v) for v in rand_eps]
    if min([float(data_beta), float(data_eps), *rand_beta, *rand_eps]) <= 0.0:
        raise RuntimeError("invalid MC-scaling objective component")
    rb, re = sum(rand_beta), sum(rand_eps)
    ratio = float(data_beta) / float(data_eps)
    f_reml = math.log(ratio / (rb / re))
    jacks = []
    for t in range(len(rand_beta)):
        jb, je = rb - rand_beta[t], re - rand_eps[t]
        jacks.append(math.log(ratio / (jb / je)) if jb > 0.0 and je > 0.0 else float("nan"))
    jacks.append(f_reml)
    as_data = tuple(math.log((rand_beta[t] / rand_eps[t]) / (rb / re)) for t in range(len(rand_beta)))
    return f_reml, tuple(jacks), as_data


def secant_search(evaluate, log_delta_at):
    """The log-delta search of ``fit_bolt_variance_components`` (``core.py:1529-1552``):
    ``evaluate(log_delta)`` returns an object with ``log_delta`` and ``f_reml``;
    ``log_delta_at(h2)`` converts a heritability. Starts at h2 = 0.25, then 0.125 (f < 0)
    or 0.5, then at most 5 secant steps clipped to [-10, 10]. Returns (best, tried)."""
    tried = []

    def ev(ld):
        r = evaluate(float(ld))
        tried.append(r)
        return r

    prev = ev(log_delta_at(0.25))
    cur = ev(log_delta_at(0.125 if prev.f_reml < 0.0 else 0.5))
    best = prev if abs(prev.f_reml) <= abs(cur.f_reml) else cur
    if abs(prev.f_reml) < abs(cur.f_reml):
        prev, cur = cur, prev
    best_from_secant = False
    for _ in range(5):
        if abs(cur.f_reml - prev.f_reml) < 1e-300:
            break
        nxt = (prev.log_delta * cur.f_reml - cur.log_delta * prev.f_reml) / (cur.f_reml - prev.f_reml)
        nxt = min(max(float(nxt), -10.0), 10.0)
        if not best_from_secant and best.log_delta == cur.log_delta and abs(nxt - cur.log_delta) < 0.01:
            break
        prev = cur
        cur = ev(nxt)
        if not best_from_secant or abs(cur.f_reml) < abs(best.f_reml):
            best = cur
            best_from_secant = True
    return best, tuple(tried)


def mc_scaling(op, ydev, G, E, *, log_delta: float, rel_tol: float, max_iter: int) -> McScaling:
    """f_REML and its parts at one log delta (``_compute_mc_scaling`` ``core.py:1421-1502``)."""
    import torch

    trials = int(G.shape[1])
    if trials + 1 > op.k:
        raise ValueError(f"{trials} trials and the data need {trials + 1} operator columns, the operator has {op.k}")
    delta = math.exp(float(log_delta))
    rhs = torch.cat([center(E * math.sqrt(delta) + G), ydev], dim=1)
    res = conj_grad(lambda P: op.apply_k(P) + delta * P, rhs, rel_tol=rel_tol, max_iter=max_iter)
    Z = res.x
    beta = score_square_sums(op, Z).cpu().numpy()
    eps = (Z * Z).sum(dim=0).cpu().numpy()
    f_reml, jacks, as_data = reml_objective(beta[trials], eps[trials], beta[:trials], eps[:trials])
    sigma2_k = float((ydev[:, 0] * Z[:, trials]).sum()) / float(max(_dim(op), 1))
    return McScaling(log_delta=float(log_delta), f_reml=f_reml, f_jacks=tuple(jacks), f_rands_as_data=as_data,
                     sigma2_k=sigma2_k, hinv_y=Z[:, trials].clone(), iterations=int(res.iterations))


def fit_variance_components(op, y, *, mc_trials: int = 15, seed: int = 0, rel_tol: float = 5e-4, max_iter: int = 500) -> VarianceFit:
    """Secant search for the log delta where f_REML = 0 (``core.py:1505-1562``)."""
    trials = max(2, int(mc_trials))
    G, E, ydev = mc_components(op, y, trials=trials, seed=seed)
    best, tried = secant_search(
        lambda ld: mc_scaling(op, ydev, G, E, log_delta=ld, rel_tol=rel_tol, max_iter=max_iter),
        lambda h2: log_delta_from_h2(op, h2))
    delta = math.exp(best.log_delta)
    return VarianceFit(log_delta=best.log_delta, delta=delta, sigma_g2=best.sigma2_k, sigma_e2=delta * best.sigma2_k,
                       h2=h2_from_log_delta(op, best.log_delta), hinv_y=best.hinv_y, evaluations=tried)


def solve_loco(op, y, *, delta: float, rel_tol: float = 5e-4, max_iter: int = 500):
    """(K_-c + delta I)^-1 P y for every GRG c (``solve_loco_hinv_y`` ``core.py:1565-1591``):
    one batched CG, column c applying the kinship of the other GRGs. Returns an
    n x (number of GRGs) FP64 device tensor."""
    import torch

    C = len(op.scales)
    if C > op.k:
        raise ValueError(f"{C} GRGs need {C} operator columns, the operator has {op.k}")
    m_per = np.asarray(op.m_per_grg, dtype=np.float64)
    m_rest = float(m_per.sum()) - m_per
    ydev = center(torch.as_tensor(np.asarray(y, dtype=np.float64).reshape(op.n, -1)[:, :1]).to(op.device))

    def apply_h(P):
        parts = op.apply_parts(P)                  # per GRG: X_g,s X_g,s^T P P, n x c
        total = sum(parts)
        cols = []
        for c in range(P.shape[1]):
            rest = (total[:, c] - parts[c][:, c]) / m_rest[c] if m_rest[c] > 0 else torch.zeros_like(total[:, c])
            cols.append(rest)
        return center(torch.stack(cols, dim=1)) + float(delta) * P

    res = conj_grad(apply_h, ydev.repeat(1, C), rel_tol=rel_tol, max_iter=max_iter)
    return res.x


__all__ = ["McScaling", "VarianceFit", "fit_variance_components", "h2_from_log_delta", "log_delta_from_h2", "mc_components",
           "mc_draws", "mc_scaling", "reml_objective", "score_square_sums", "secant_search", "solve_loco"]
