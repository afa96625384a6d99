"""Standardized X^T X operator (``pygrgl_spmv.pca``) and its row-weighted sums.

Counterpart in the reference: ``scripts/pca_bench/__main__.py`` (``PreparedXTX``
``:184-420``, ``row_weighted_sum`` ``:32-61``). The operator is checked against
the dense Z^T Z built from the CPU oracle's genotype and missingness matrices
(FP64 within 1e-10 normwise: the centring cancels leading digits); the leading eigenvalues against ``numpy.linalg.eigvalsh``
of the dense matrix; the reduction against numpy sums.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import grg_oracle as O
from tests._builders import build_hip_layout, write_synthetic_artifact

pytestmark = pytest.mark.gpu


def _cfg(seed, missing):
    from pygrgl_spmv.synth import SynthConfig

    return SynthConfig(num_samples=200, num_mutations=900, seed=seed, ploidy=2, num_trees=80, keep_fraction=0.3,
                       window_fraction=0.1, num_segments=2, missing_fraction=missing, missing_pool=4)


def _dense_z(g):
    """Z = (X + Miss diag(f) - 1 c^T) diag(1/s) from the oracle's sweeps (identity inputs)."""
    I, M = g.num_individuals, g.num_mutations
    eye = np.eye(I)
    miss = np.zeros((I, M))
    X = O.matmul(g, eye, "up", by_individual=True, miss=miss)          # (I, M) dosages, miss += Miss
    ones = np.ones((1, I))
    counts = O.matmul(g, ones, "up", by_individual=True)[0]
    mcount = np.zeros((1, M))
    O.matmul(g, ones, "up", by_individual=True, miss=mcount)
    denom = float(g.num_samples) - mcount[0]
    f = np.where(denom != 0.0, counts / np.where(denom != 0.0, denom, 1.0), 0.0)
    c = g.ploidy * f
    s = np.sqrt(g.ploidy * f * (1.0 - f))
    s = np.where(s == 0.0, 1.0, s)
    return (X + miss * f[None, :] - c[None, :]) / s[None, :]


@pytest.mark.parametrize("missing", [0.0, 0.05])
def test_xtx_operator_matches_dense(artifact_dir, missing):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.pca import PreparedXTX, xtx_top_eigenpairs
    from tests._builders import full_requirements

    art = write_synthetic_artifact(artifact_dir, f"pca-{missing}", _cfg(701, missing))
    g = O.load_artifact(art)
    Z = _dense_z(g)
    A = Z.T @ Z
    layout = build_hip_layout([art], requirements=full_requirements(max_k_up=16, max_k_down=16))
    rng = np.random.default_rng(7100)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        assert grg.has_missing_data == (missing > 0)
        with PreparedXTX(grg, k=16) as op:
            for q in (16, 5, 1):
                V = rng.standard_normal((g.num_mutations, q))
                got = op.apply(torch.from_numpy(V).to(op.device)).cpu().numpy()
                # the centring subtracts the mean dosage term from X a and from X^T u (as the
                # reference's steps do), which cancels a few leading digits: 1e-10, not 1e-12
                assert O.normwise_rel_err(got, A @ V) <= 1e-10
            lam, U, res = xtx_top_eigenpairs(op, 4, oversample=12, iters=80)
            top = np.sort(np.linalg.eigvalsh(A))[::-1][:4]
            np.testing.assert_allclose(lam.cpu().numpy(), top, rtol=1e-6)
            with pytest.raises(ValueError, match="exceed"):
                op.apply(torch.zeros((g.num_mutations, 17), dtype=torch.float64, device=op.device))


@pytest.mark.parametrize("rows,cols,ld,blocks", [(1, 1, 1, 1), (16, 5000, 16, 512), (7, 12345, 9, 64), (256, 300, 256, 3)])
def test_row_weighted_sum(rows, cols, ld, blocks):
    import torch

    from pygrgl_spmv.backends.hip import ffi

    lib = ffi.hip_library()
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(rows * 7 + cols)
    x = rng.standard_normal(max(1, (cols - 1) * ld + rows))
    w = rng.standard_normal(cols)
    want = np.array([sum(x[j + m * ld] * w[m] for m in range(cols)) if cols < 50 else
                     float(np.dot(x[j + np.arange(cols) * ld], w)) for j in range(rows)])
    for dt, code, tol in ((torch.float64, ffi.F64, 1e-12), (torch.float32, ffi.F32C, 1e-5)):
        xt = torch.from_numpy(x).to(dev).to(dt)
        wt = torch.from_numpy(w).to(dev).to(dt)
        out = torch.empty(rows, dtype=dt, device=dev)
        part = torch.empty(blocks * rows, dtype=torch.float64, device=dev)
        s = torch.cuda.current_stream(dev).cuda_stream
        lib.row_weighted_sum(code, xt.data_ptr(), rows, cols, ld, wt.data_ptr(), out.data_ptr(), part.data_ptr(), blocks, s)
        first = out.cpu().numpy().astype(np.float64)
        lib.row_weighted_sum(code, xt.data_ptr(), rows, cols, ld, wt.data_ptr(), out.data_ptr(), part.data_ptr(), blocks, s)
        np.testing.assert_array_equal(out.cpu().numpy().astype(np.float64), first)  # deterministic
        ref = want if dt == torch.float64 else np.array([float(np.dot(x[j + np.arange(cols) * ld].astype(np.float32).astype(np.float64),
                                                                       w.astype(np.float32).astype(np.float64))) for j in range(rows)])
        assert np.max(np.abs(first - ref)) <= tol * max(1.0, np.max(np.abs(ref)))
    with pytest.raises(RuntimeError, match="rows must be"):
        lib.row_weighted_sum(ffi.F64, 0, 300, 1, 300, 0, 0, 0, 1, 0)
