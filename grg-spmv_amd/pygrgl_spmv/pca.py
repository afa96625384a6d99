"""Standardized X^T X operator on one GRG, and its leading eigenpairs.

Restates the prepared operator of the reference's PCA benchmark
(``scripts/pca_bench/__main__.py``: ``PreparedXTX.initialize`` ``:200-239``,
``_body`` ``:334-358``, ``row_weighted_sum`` ``:32-61``) on the HIP backend.

With X the individuals x mutations dosage matrix, Miss the missingness
indicator and f the allele frequencies (allele count over the non-missing
haplotype count, from one up sweep of ones), missing genotypes are imputed as
X + Miss diag(f); with centre c = ploidy f and scale s = sqrt(ploidy f (1 - f))
(1 where that is 0), Z = (X + Miss diag(f) - 1 c^T) diag(1 / s) and the operator
is V -> Z^T Z V (mutations x q). One application is one down sweep and one up
sweep by individual (with the missing-data input / output when the GRG has
missing data) and these O(M q + I q) steps in FP64 on the device:

1. a = V^T diag(1/s); the down sweep's missing-data input is a diag(f);
2. d_j = sum_m a[j, m] c[m] (``grg_hip_row_weighted_sum``);
3. u = down(a) - d 1^T, which is Z V; e_j = sum_i u[j, i];
4. up(u) (plus the missing-data output times f), times 1/s, minus e c/s.

The sweeps run in the runtime's value type; the operator's vectors are FP64.
"""

from __future__ import annotations

import numpy as np


class PreparedXTX:
    """Z^T Z on one GRG of an entered ``HipRuntime`` (runtime requirements: up
    missing-data output when the GRG has missing data; max_k_up / max_k_down >= k),
    applied to up to ``k`` columns at a time."""

    def __init__(self, grg, *, k: int, n_blocks: int = 512):
        import torch

        from pygrgl_spmv.backends.hip import ffi

        self.grg = grg
        self.k = int(k)
        if not 1 <= self.k <= 256:
            raise ValueError("k must be in [1, 256]")
        self.has_missing = bool(grg.has_missing_data)
        self.shape = (int(grg.num_mutations), int(grg.num_mutations))
        self._lib = ffi.hip_library()
        self._dtype_code = ffi.F64 if np.dtype(grg._runtime.layout.dtype) == np.float64 else ffi.F32C
        self._n_blocks = int(n_blocks)
        self._entered = []
        self.down = self._enter(grg.prepare_matmul_cuda(direction="down", k=self.k, by_individual=True, use_miss=self.has_missing))
        self.up = self._enter(grg.prepare_matmul_cuda(direction="up", k=self.k, by_individual=True, use_miss=self.has_missing))
        dev = self.down.input.device
        self.device = dev
        self._tdt = self.down.input.dtype
        self._partials = torch.empty((self._n_blocks * self.k,), dtype=torch.float64, device=dev)
        self._d = torch.empty((self.k,), dtype=self._tdt, device=dev)
        self._initialize()

    def _enter(self, manager):
        op = manager.__enter__()
        self._entered.append(manager)
        return op

    def close(self) -> None:
        for m in reversed(self._entered):
            m.__exit__(None, None, None)
        self._entered = []

    def __enter__(self) -> "PreparedXTX":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def _initialize(self) -> None:
        """Allele frequencies, centre and scale (reference ``:200-239``): one up sweep
        of ones over the individuals, with the missing-data output."""
        import torch

        up = self.up
        up.input.fill_(1.0)
        if self.has_missing:
            up.miss_output.fill_(0.0)
        up()
        counts = up.output[0].to(torch.float64).clone()
        missing = up.miss_output[0].to(torch.float64).clone() if self.has_missing else torch.zeros_like(counts)
        denom = float(self.grg.num_samples) - missing
        valid = denom != 0.0
        f = torch.where(valid, counts / torch.where(valid, denom, torch.ones_like(denom)), torch.zeros_like(counts))
        ploidy = float(self.grg.ploidy)
        self.freqs = f
        self.center = ploidy * f
        sigma = torch.sqrt(ploidy * f * (1.0 - f))
        self.sigma = torch.where(sigma == 0.0, torch.ones_like(sigma), sigma)
        self.inv_sigma = 1.0 / self.sigma
        self.center_over_sigma = self.center * self.inv_sigma
        self._center_t = self.center.to(self._tdt)

    def apply(self, V):
        """Z^T Z V for an (M, q) FP64 device tensor V, q <= k; returns (M, q) FP64."""
        import torch

        if V.dim() != 2 or int(V.shape[0]) != self.shape[1]:
            raise ValueError(f"X^T X operand has shape {tuple(V.shape)}, expected ({self.shape[1]}, q)")
        q = int(V.shape[1])
        if not 1 <= q <= self.k:
            raise ValueError(f"q = {q} columns exceed the prepared width k = {self.k}")
        down, up = self.down, self.up
        a = (V.T.to(torch.float64) * self.inv_sigma[None, :])          # (q, M)
        down.input.zero_()
        down.input[:q].copy_(a)
        if self.has_missing:
            down.miss_input.zero_()
            down.miss_input[:q].copy_(a * self.freqs[None, :])
        # d_j = sum_m a[j, m] c[m], on the down input view (columns of the arena, ld = its row stride)
        x = down.input
        self._lib.row_weighted_sum(self._dtype_code, x.data_ptr(), self.k, self.shape[1], x.stride(1), self._center_t.data_ptr(),
                                   self._d.data_ptr(), self._partials.data_ptr(), self._n_blocks,
                                   torch.cuda.current_stream(self.device).cuda_stream)
        d = self._d[:q].to(torch.float64)
        down()
        u = down.output[:q].to(torch.float64) - d[:, None]                # (q, I) = (Z V)^T
        e = u.sum(dim=1)
        up.input.zero_()
        up.input[:q].copy_(u)
        if self.has_missing:
            up.miss_output.zero_()
        up()
        out = up.output[:q].to(torch.float64)
        if self.has_missing:
            out = out + up.miss_output[:q].to(torch.float64) * self.freqs[None, :]
        out = out * self.inv_sigma[None, :] - e[:, None] * self.center_over_sigma[None, :]
        return out.T.contiguous()


def xtx_top_eigenpairs(op: PreparedXTX, r: int, *, oversample: int = 8, iters: int = 20, seed: int = 0):
    """Leading eigenpairs of Z^T Z by block subspace iteration: each iteration is one
    operator application to an M x q block and a Householder QR in FP64 on the GPU;
    a Rayleigh-Ritz step on the final block gives (values descending, vectors M x r,
    residual norms)."""
    import torch

    q = min(int(r) + int(oversample), op.k)
    if q < int(r):
        raise ValueError(f"r={r} exceeds the operator's column capacity {op.k}")
    rng = np.random.default_rng(seed)
    Q, _ = torch.linalg.qr(torch.from_numpy(rng.standard_normal((op.shape[1], q))).to(op.device))
    for _ in range(int(iters)):
        Q, _ = torch.linalg.qr(op.apply(Q))
    AQ = op.apply(Q)
    T = Q.T @ AQ
    T = 0.5 * (T + T.T)
    vals, vecs = torch.linalg.eigh(T)
    order = torch.argsort(vals, descending=True)[: int(r)]
    U = Q @ vecs[:, order]
    lam = vals[order]
    res = torch.linalg.norm(AQ @ vecs[:, order] - U * lam, dim=0)
    return lam, U, res


__all__ = ["PreparedXTX", "xtx_top_eigenpairs"]
