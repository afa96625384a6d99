"""Learned CG warm start for (K + delta I) x = P y.

No reference counterpart: the reference's solver starts from zero
(``scripts/bolt_lmm_inf/core.py:1322``). The solution is linear in y, so
everything is normalised by ||P y||:

* inputs (r + 1 per phenotype): the projections U^T y_hat of y_hat = P y / ||P y||
  onto the top-r eigenvectors U of K, and log(delta);
* outputs (r + 1): coefficients c and a such that
  x0 = ||P y|| (U c + a * q),  q = y_hat - U U^T y_hat
  (the part of y_hat outside the principal subspace);
* training targets come from solves of the plain method: for a solution x*,
  c = U^T x* / ||P y|| and a = <x* / ||P y|| - U c, q> / ||q||^2, the best
  such coefficients in the least-squares sense.

``closed_form_coefficients`` is the same start built from the eigenvalues
alone (c_i = (u_i^T y_hat) / (lambda_i + delta), a = 1 / (lambda_rest + delta) with
lambda_rest the mean of K's remaining eigenvalues). The network
(``mlp.DenseNet``, HIP kernels) learns the correction from these closed-form
coefficients to the best ones; the closed form alone is also reported as a
second baseline.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from pygrgl_spmv.lmm.mlp import DenseNet
from pygrgl_spmv.lmm.operator import center


@dataclass
class Basis:
    U: object      # n x r FP64 device tensor
    lam: object    # r FP64 device tensor
    trace: float   # trace(K)


def _normalised(basis: Basis, y):
    import torch

    Py = center(y.to(torch.float64))
    norm = torch.linalg.norm(Py, dim=0)
    yh = Py / norm
    proj = basis.U.T @ yh                 # r x c
    q = yh - basis.U @ proj               # n x c
    return norm, yh, proj, q


def features(basis: Basis, y, delta):
    """(c x (r + 1)) FP32 network inputs."""
    import torch

    _, _, proj, _ = _normalised(basis, y)
    f = torch.cat([proj.T, torch.log(delta).reshape(-1, 1)], dim=1)
    return f.to(torch.float32)


def targets(basis: Basis, y, x_star):
    """(c x (r + 1)) FP32 regression targets from solutions x* of the plain method."""
    import torch

    norm, _, _, q = _normalised(basis, y)
    xh = x_star.to(torch.float64) / norm
    c = basis.U.T @ xh
    rest = xh - basis.U @ c
    a = (rest * q).sum(dim=0) / (q * q).sum(dim=0)
    return torch.cat([c.T, a.reshape(-1, 1)], dim=1).to(torch.float32)


def start_from_coefficients(basis: Basis, y, coef):
    """x0 = ||P y|| (U c + a q) for coefficient rows ``coef`` (c x (r + 1))."""
    import torch

    norm, _, _, q = _normalised(basis, y)
    coef = coef.to(torch.float64)
    r = basis.U.shape[1]
    x = basis.U @ coef[:, :r].T + q * coef[:, r]
    return x * norm


def closed_form_coefficients(basis: Basis, y, delta):
    import torch

    _, _, proj, _ = _normalised(basis, y)
    n = basis.U.shape[0]
    r = basis.U.shape[1]
    lam_rest = (basis.trace - float(basis.lam.sum())) / max(1, n - 1 - r)
    c = proj / (basis.lam.reshape(-1, 1) + delta.reshape(1, -1))
    a = 1.0 / (lam_rest + delta)
    return torch.cat([c.T, a.reshape(-1, 1)], dim=1)


class WarmStartNet:
    """The network predicts a correction to the closed-form coefficients: its output
    (de-standardised with training-set statistics) is added to
    ``closed_form_coefficients``. ``DenseNet([r + 1, *hidden, r + 1])``, trained with
    Adam on the GPU.

    The network is only deployed when it earns it: ``choose`` takes the mean CG
    iterations of held-out validation solves from both starts, and ``start`` uses
    the network only if it needed fewer iterations there by at least ``margin``
    (otherwise the closed-form start)."""

    def __init__(self, basis: Basis, *, hidden=(256, 256), seed: int = 0):
        self.basis = basis
        r = int(basis.U.shape[1])
        self.net = DenseNet([r + 1, *hidden, r + 1], device=basis.U.device, seed=seed)
        self.stats = None
        self.use_network = True
        self.validation = None

    def choose(self, network_iterations: float, closed_form_iterations: float, *, margin: float = 0.5) -> bool:
        """Deploy the network only if its validation starts needed fewer CG iterations than the
        closed-form starts by at least ``margin`` (mean per phenotype). Returns the choice."""
        self.validation = {"network_mean_iterations": float(network_iterations),
                           "closed_form_mean_iterations": float(closed_form_iterations), "margin": float(margin)}
        self.use_network = float(network_iterations) <= float(closed_form_iterations) - float(margin)
        self.validation["network_deployed"] = self.use_network
        return self.use_network

    def fit(self, y, delta, x_star, *, epochs: int = 300, batch: int = 128, lr: float = 1e-3, seed: int = 0):
        """Train on solutions ``x_star`` of the plain method for phenotypes ``y`` (n x c)
        with shifts ``delta`` (c); returns the per-epoch training loss."""
        import torch

        F = features(self.basis, y, delta)
        T = targets(self.basis, y, x_star) - closed_form_coefficients(self.basis, y, delta).to(torch.float32)
        fm, fs = F.mean(dim=0), F.std(dim=0) + 1e-6
        tm, ts = T.mean(dim=0), T.std(dim=0) + 1e-12
        self.stats = (fm, fs, tm, ts)
        return self.net.train((F - fm) / fs, (T - tm) / ts, epochs=epochs, batch=batch, lr=lr, seed=seed)

    def coefficients(self, y, delta):
        import torch

        fm, fs, tm, ts = self.stats
        F = features(self.basis, y, delta)
        correction = self.net.forward((F - fm) / fs) * ts + tm
        return closed_form_coefficients(self.basis, y, delta) + correction.to(torch.float64)

    def network_start(self, y, delta):
        return start_from_coefficients(self.basis, y, self.coefficients(y, delta))

    def start(self, y, delta):
        """The deployed start: the network's, or the closed form when validation rejected it."""
        if not self.use_network:
            return start_from_coefficients(self.basis, y, closed_form_coefficients(self.basis, y, delta))
        return self.network_start(y, delta)


__all__ = ["Basis", "WarmStartNet", "closed_form_coefficients", "features", "start_from_coefficients", "targets"]
