"""Standardized-genotype kinship operator over the GRGs of a ``HipRuntime``.

Restates the GRG operator of the reference's BOLT-LMM-inf script with the
all-ones covariate basis (no covariates): ``scripts/bolt_lmm_inf/core.py``

* per-variant statistics ``attach_bed_stats`` ``:935-967``: mean = sum_g / n,
  mean-centred norm^2 = sumsq_g - sum_g^2 / n, scale = sqrt((n - 1) / norm^2)
  (0 for a monomorphic variant, which leaves the model);
* ``scores`` ``:1241-1252`` (X_s^T P v), ``apply_x`` ``:1254-1266`` (P X_s w) and
  ``apply_k`` ``:1282-1301``: K v = P (1 / M) sum_chrom X_s X_s^T P v, with P the
  projection off the covariate basis (here: centring).

With P = centring the mean corrections of ``scores`` and ``apply_x`` vanish
(P v sums to zero; P removes the constant of ``apply_x``), so one application
is: centre, up sweep by individual with the per-variant scale^2 applied inside
the write-out launch, down sweep by individual (its input is the up output in
the shared arena, no copy), accumulate over GRGs, divide by M, centre. The
sweeps run in the runtime's dtype (FP32-compensated or FP64); the
accumulation across GRGs, the division and the centring stay in FP64 on the
GPU.
"""

from __future__ import annotations

import numpy as np


def center(V):
    """P V for the all-ones covariate basis: subtract each column's mean (FP64 on the device)."""
    return V - V.mean(dim=0, keepdim=True)


class StandardizedGrgOperator:
    """K = P X_s X_s^T P / M over every GRG of an entered ``HipRuntime``, applied to
    up to ``k`` columns at a time (n individuals x k, FP64 device tensors)."""

    def __init__(self, runtime, *, k: int):
        import torch

        self.runtime = runtime
        self.k = int(k)
        self.device = runtime.device
        grgs = runtime.grgs
        if not grgs:
            raise ValueError("the runtime holds no GRG")
        n = int(grgs[0].num_individuals)
        if any(int(g.num_individuals) != n for g in grgs):
            raise ValueError("every GRG must cover the same individuals")
        self.n = n
        tdt = runtime._torch_dtype()
        self._ops = []
        self.means, self.scales = [], []
        self.m_per_grg = []
        self._inv_scales = []
        self.m_proj = 0
        for grg in grgs:
            ones = np.ones((1, grg.num_samples))
            sum_g = grg.matmul(ones, "up")[0].astype(np.float64)               # sum over individuals of dosage
            sumsq_g = grg.matmul(ones, "up", init="xtx")[0].astype(np.float64)  # sum of dosage^2
            mean = sum_g / n
            norm2 = sumsq_g - sum_g * sum_g / n
            scale = np.where(norm2 > 0.0, np.sqrt((n - 1.0) / np.where(norm2 > 0.0, norm2, 1.0)), 0.0)
            up = grg.prepare_matmul_cuda(direction="up", k=self.k, by_individual=True)
            down = grg.prepare_matmul_cuda(direction="down", k=self.k, by_individual=True)
            up.set_output_scale(torch.from_numpy(scale * scale).to(device=self.device, dtype=tdt))
            self._ops.append((up, down))
            self.means.append(mean)
            self.scales.append(scale)
            inv = np.where(scale > 0.0, 1.0 / np.where(scale > 0.0, scale, 1.0), 0.0)
            self._inv_scales.append(torch.from_numpy(inv).to(self.device))
            self.m_per_grg.append(int(np.count_nonzero(scale > 0.0)))
            self.m_proj += self.m_per_grg[-1]
        if self.m_proj == 0:
            raise ValueError("no polymorphic variant: the operator is empty")
        # sum over model variants of ||P X_s[:, j]||^2 (= n - 1 each, core.py:950-954)
        self.xfro2 = float(self.m_proj) * (n - 1.0)

    def _pad(self, V):
        import torch

        if V.shape[1] > self.k:
            raise ValueError(f"at most {self.k} columns per application, got {V.shape[1]}")
        if V.shape[1] == self.k:
            return V
        return torch.cat([V, torch.zeros((V.shape[0], self.k - V.shape[1]), dtype=V.dtype, device=V.device)], dim=1)

    def apply_k(self, V):
        """K V for V (n x c, c <= k) as an FP64 device tensor."""
        import torch

        c = int(V.shape[1])
        W = self._pad(center(V.to(dtype=torch.float64)))
        out = torch.zeros((self.n, self.k), dtype=torch.float64, device=self.device)
        for up, down in self._ops:
            up.input.copy_(W.T)
            up()
            down()  # its input is up's output in the shared arena
            out += down.output.T.to(torch.float64)
        out /= float(self.m_proj)
        return center(out)[:, :c]

    def apply_parts(self, V):
        """Per GRG g, X_g,s X_g,s^T P V (n x c FP64, neither divided by M nor centred):
        the terms whose sum over the included GRGs gives ``apply_k`` with a GRG left out
        (``apply_k(..., exclude_chrom=...)``, reference ``core.py:1282-1301``)."""
        import torch

        c = int(V.shape[1])
        W = self._pad(center(V.to(dtype=torch.float64)))
        parts = []
        for up, down in self._ops:
            up.input.copy_(W.T)
            up()
            down()
            parts.append(down.output.T[:, :c].to(torch.float64).clone())
        return parts

    def scores(self, V):
        """Per GRG g, X_g,s^T P V (M_g x c FP64; 0 on monomorphic variants), reference
        ``scores`` ``core.py:1241-1252``: one up sweep, whose write-out applies scale^2,
        then a division by the scale in FP64."""
        import torch

        c = int(V.shape[1])
        W = self._pad(center(V.to(dtype=torch.float64)))
        out = []
        for (up, _), inv in zip(self._ops, self._inv_scales):
            up.input.copy_(W.T)
            up()
            out.append(up.output.T[:, :c].to(torch.float64) * inv[:, None])
        return out

    def genetic_values(self, betas):
        """P X_s beta for per-GRG effect vectors ``betas[g]`` (M_g x c numpy), as n x c FP64."""
        import torch

        c = int(betas[0].shape[1])
        if c > self.k:
            raise ValueError(f"at most {self.k} columns, got {c}")
        out = torch.zeros((self.n, self.k), dtype=torch.float64, device=self.device)
        for (_, down), scale, beta in zip(self._ops, self.scales, betas):
            w = np.zeros((beta.shape[0], self.k))
            w[:, :c] = beta * scale[:, None]
            down.input.copy_(torch.from_numpy(w.T.copy()).to(self.device))
            down()
            out += down.output.T.to(torch.float64)
        return center(out)[:, :c]


__all__ = ["StandardizedGrgOperator", "center"]
