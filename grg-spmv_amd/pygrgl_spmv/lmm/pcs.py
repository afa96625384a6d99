"""Top principal components of the kinship operator K by block subspace iteration.

Each iteration applies K to an n x q block through the GRG sweeps and
re-orthonormalises the block with a Householder QR in FP64 on the GPU; a
Rayleigh-Ritz step on the final block gives the leading eigenpairs. No
reference counterpart on the hot path: the reference's PCA script
(``scripts/pca_bench/__main__.py``) is out of scope; the eigenvectors here only
feed the CG warm start.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from pygrgl_spmv.lmm.operator import center


@dataclass
class Eigenpairs:
    vectors: object   # n x r FP64 device tensor, orthonormal, centred
    values: object    # r FP64 device tensor, descending
    residuals: object  # r, ||K u - lambda u|| per pair


def top_eigenpairs(op, r: int, *, oversample: int = 12, iters: int = 15, seed: int = 0) -> Eigenpairs:
    import torch

    q = min(int(r) + int(oversample), int(op.k))
    if q < int(r):
        raise ValueError(f"r={r} exceeds the operator's column capacity {op.k}")
    rng = np.random.default_rng(seed)
    Q = center(torch.from_numpy(rng.standard_normal((op.n, q))).to(op.device))
    Q, _ = torch.linalg.qr(Q)
    for _ in range(int(iters)):
        Q, _ = torch.linalg.qr(center(op.apply_k(Q)))
    KQ = op.apply_k(Q)
    T = Q.T @ KQ
    T = 0.5 * (T + T.T)
    vals, vecs = torch.linalg.eigh(T)
    order = torch.argsort(vals, descending=True)[: int(r)]
    U = Q @ vecs[:, order]
    lam = vals[order]
    KU = KQ @ vecs[:, order]
    res = torch.linalg.norm(KU - U * lam, dim=0)
    return Eigenpairs(vectors=U, values=lam, residuals=res)


__all__ = ["Eigenpairs", "top_eigenpairs"]
