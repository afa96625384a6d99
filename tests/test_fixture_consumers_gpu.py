"""The consumers of the HIP sweeps on the reference's own 200-sample GRG
(``tests/golden/grgl``, copied from ``pygrgl_spmv/tests/data``; it has missing
genotypes), against dense matrices built from the IGD genotype file of the same
dataset: the captured adaptor replays (every captured configuration), the
standardized X^T X operator with missing-genotype imputation and its leading
eigenvalues, and the kinship operator with the variance-component search. FP64:
1e-12 normwise for the sweeps, 1e-10 where centring cancels digits, 1e-9 for
quantities behind a CG solve."""

from __future__ import annotations

import contextlib
from pathlib import Path

import numpy as np
import pytest

from oracle import grg_oracle as O
from oracle import lmm_oracle as L
from pygrgl_spmv import convert
from pygrgl_spmv.grg.grgl_file import read_grg
from tests._builders import build_hip_layout, full_requirements
from tests._igd import genotype_columns, read_igd

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "grgl"
T200 = GOLDEN / "test-200-samples.miss.final.grg"
IGD = GOLDEN / "test-200-samples.miss.igd"


def _dip(H):
    D = np.zeros((H.shape[0] // 2, H.shape[1]))
    np.add.at(D, np.arange(H.shape[0]) // 2, H)
    return D


@pytest.fixture(scope="module")
def t200(tmp_path_factory):
    g = read_grg(T200)
    X, Miss = genotype_columns(read_igd(IGD), g.positions, g.ref_alleles, g.alleles)
    return convert(T200, tmp_path_factory.mktemp("fixture"), name="t200"), g, X, Miss


def test_captured_replays_on_fixture(t200):
    from pygrgl_spmv.adaptor import CaptureSpec, RunConfigs, load_grg_spmv_single, make_backend_hip
    from pygrgl_spmv.grg import RuntimeRequirements

    art, g, X, Miss = t200
    D, Dm = _dip(X), _dip(Miss)
    xtx_bias = (D * D).sum(axis=0) - X.sum(axis=0)     # init='xtx' adds 2 x (individuals with both copies)
    k = 4
    req = RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=True, need_up_miss_output=True,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=True)
    ops = (CaptureSpec("up", k), CaptureSpec("down", k), CaptureSpec("down", k, use_miss=True),
           CaptureSpec("up", k, use_miss=True), CaptureSpec("up", k, by_individual=True, init_mode="xtx"),
           CaptureSpec("down", k, by_individual=True))
    rng = np.random.default_rng(91)
    with contextlib.ExitStack() as stack:
        grg = load_grg_spmv_single(art, make_backend_hip(capture=True), RunConfigs(req=req, capture_ops=ops), stack)
        for rep in range(2):
            for kk in (k, 3):
                x = rng.standard_normal((kk, g.num_samples))
                y = rng.standard_normal((kk, g.num_mutations))
                m = rng.standard_normal((kk, g.num_mutations))
                xi = rng.standard_normal((kk, g.num_individuals))
                assert O.normwise_rel_err(grg.matmul(x, "up"), x @ X) <= 1e-12
                assert O.normwise_rel_err(grg.matmul(y, "down"), y @ X.T) <= 1e-12
                assert O.normwise_rel_err(grg.matmul(y, "down", miss=m), y @ X.T + m @ Miss.T) <= 1e-12
                acc = np.zeros((kk, g.num_mutations))
                assert O.normwise_rel_err(grg.matmul(x, "up", miss=acc), x @ X) <= 1e-12
                assert O.normwise_rel_err(acc, x @ Miss) <= 1e-12
                got = grg.matmul(xi, "up", by_individual=True, init="xtx")
                assert O.normwise_rel_err(got, xi @ D + xtx_bias[None, :]) <= 1e-12
                assert O.normwise_rel_err(grg.matmul(y, "down", by_individual=True), y @ D.T) <= 1e-12
        ones = np.ones((2, g.num_individuals))
        np.testing.assert_array_equal(grg.matmul(ones, "up", by_individual=True, init="xtx"),
                                      np.tile((D * D).sum(axis=0), (2, 1)))     # sum of dosage^2
        assert Dm.sum() > 0


def test_xtx_operator_with_imputation_on_fixture(t200):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.pca import PreparedXTX, xtx_top_eigenpairs

    art, g, X, Miss = t200
    S = g.num_samples
    counts, mcount = X.sum(axis=0), Miss.sum(axis=0)
    denom = S - mcount
    f = np.where(denom != 0, counts / np.where(denom != 0, denom, 1), 0.0)
    s = np.sqrt(2.0 * f * (1.0 - f))
    s = np.where(s == 0.0, 1.0, s)
    Z = (_dip(X) + _dip(Miss) * f[None, :] - 2.0 * f[None, :]) / s[None, :]
    rng = np.random.default_rng(92)
    with HipRuntime(build_hip_layout([art], requirements=full_requirements(max_k_up=16, max_k_down=16))) as rt:
        (grg,) = rt.grgs
        assert grg.has_missing_data
        with PreparedXTX(grg, k=16) as op:
            np.testing.assert_allclose(op.freqs.cpu().numpy(), f, rtol=1e-14, atol=1e-15)
            for q in (16, 3):
                V = rng.standard_normal((g.num_mutations, q))
                got = op.apply(torch.from_numpy(V).to(op.device)).cpu().numpy()
                exp = Z.T @ (Z @ V)
                assert O.normwise_rel_err(got.T, exp.T) <= 1e-10        # per column of V
                zero = ~Z.any(axis=0)                                    # mutations without carriers
                assert zero.sum() > 0 and np.abs(got[zero]).max() <= 1e-10 * np.abs(exp).max()
            lam, _, _ = xtx_top_eigenpairs(op, 4, oversample=12, iters=80)
    top = np.sort(np.linalg.eigvalsh(Z @ Z.T))[::-1][:4]       # the nonzero spectrum of Z^T Z
    np.testing.assert_allclose(lam.cpu().numpy(), top, rtol=1e-6)


def test_kinship_and_variance_search_on_fixture(t200):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.lmm import StandardizedGrgOperator
    from pygrgl_spmv.lmm import variance as V
    from pygrgl_spmv.grg import RuntimeRequirements

    art, g, X, _ = t200
    (Xs, keep), = L.standardized([_dip(X)])
    rng = np.random.default_rng(93)
    y = Xs @ rng.standard_normal(Xs.shape[1]) * np.sqrt(0.5 / Xs.shape[1]) + rng.standard_normal(Xs.shape[0]) * np.sqrt(0.5)
    req = RuntimeRequirements(max_k_up=8, max_k_down=8, need_down_miss_input=False, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=True)
    with HipRuntime(build_hip_layout([art], requirements=req)) as rt:
        op = StandardizedGrgOperator(rt, k=8)
        assert op.m_proj == int(keep.sum())
        Vm = rng.standard_normal((op.n, 5))
        got = op.apply_k(torch.from_numpy(Vm).to(op.device)).cpu().numpy()
        assert O.normwise_rel_err(got.T, (L.kinship_of([Xs]) @ Vm).T) <= 1e-12
        fit = V.fit_variance_components(op, y, mc_trials=5, seed=8, rel_tol=1e-12, max_iter=1000)
        betas, noise = V.mc_draws(op.scales, op.n, op.m_proj, trials=5, seed=8)
    G = L.genetic_values([Xs], [keep], betas)
    for ev in fit.evaluations:
        exp = L.mc_scaling([Xs], G, noise, y, ev.log_delta)
        assert ev.f_reml == pytest.approx(exp["f_reml"], abs=1e-9)
        assert ev.sigma2_k == pytest.approx(exp["sigma2_k"], rel=1e-9)
    assert 0.0 < fit.h2 < 1.0
