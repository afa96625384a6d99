"""Batched conjugate gradients for (K + delta I) x = P y on the GPU, in FP64.

Follows the reference's ``bolt_conj_grad_solve`` (``scripts/bolt_lmm_inf/core.py:1304-1353``):
full-batch CG over all right-hand sides together, no active-column mask, no
denominator guard, the residual re-projected after every update, and every
column iterated until all columns reach the tolerance (or ``max_iter``).

Two additions, both off by default so the zero-start solve is the reference's:

* ``x0``: a starting point (the learned warm start). The residual starts at
  P y - H x0, one extra operator application.
* the stopping test compares ||r|| with ||P y|| (the zero-start initial
  residual), so a warm start is judged at the same residual as the zero start.

Each column's first iteration below the tolerance is recorded, so the gain of
a warm start can be counted per right-hand side.
"""

from __future__ import annotations

from dataclasses import dataclass

from pygrgl_spmv.lmm.operator import center


@dataclass
class CgResult:
    x: object            # n x c FP64 device tensor (projected)
    iterations: int      # iterations of the batch
    column_iterations: object  # per column, first iteration with ||r|| <= tol ||P y|| (-1 if never)
    rel_residual: object  # per column, final ||r|| / ||P y||
    applications: int    # operator applications, including the warm start's initial residual


def conj_grad(apply_h, rhs, *, x0=None, rel_tol: float = 1e-6, max_iter: int = 500, project=center) -> CgResult:
    """Solve H X = P rhs column by column (all columns iterated together).

    ``apply_h(P)`` returns H P for an n x c FP64 device tensor. ``rhs`` is n x c.
    """
    import torch

    B = project(rhs.to(torch.float64))
    b2 = (B * B).sum(dim=0)
    applications = 0
    if x0 is None:
        X = torch.zeros_like(B)
        R = B.clone()
    else:
        X = project(x0.to(torch.float64)).clone()
        R = project(B - apply_h(X))
        applications += 1
    Pd = R.clone()
    r2 = (R * R).sum(dim=0)
    tol2 = float(rel_tol) ** 2
    col_it = torch.full((B.shape[1],), -1, dtype=torch.int64, device=B.device)
    col_it[r2 <= tol2 * b2] = 0
    it = 0
    for it in range(1, int(max_iter) + 1):
        HP = apply_h(Pd)
        applications += 1
        alpha = r2 / (Pd * HP).sum(dim=0)
        X += Pd * alpha
        R -= HP * alpha
        R = project(R)
        r2_new = (R * R).sum(dim=0)
        done = r2_new <= tol2 * b2
        col_it[done & (col_it < 0)] = it
        if bool(done.all()):
            r2 = r2_new
            break
        beta = r2_new / r2
        Pd = R + Pd * beta
        r2 = r2_new
    return CgResult(x=project(X), iterations=it, column_iterations=col_it, rel_residual=torch.sqrt(r2 / b2),
                    applications=applications)


def shifted(op, delta):
    """H = K + diag-per-column delta: apply_h(P) = K P + P * delta (delta: c FP64 device tensor)."""

    def apply_h(P):
        return op.apply_k(P) + P * delta[: P.shape[1]]

    return apply_h


__all__ = ["CgResult", "conj_grad", "shifted"]
