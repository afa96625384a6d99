"""Mixed-model consumer on the GPU: kinship operator, CG, principal components,
network kernels and the learned warm start, each against its CPU oracle.

Operator: FP64 sweeps within 1e-12 of the dense oracle K = P X_s X_s^T P / M,
FP32-compensated sweeps within 1e-5. CG: solutions within 1e-6 of a dense direct
solve. Network kernels: loss, gradients and Adam steps within 1e-4 relative of
the FP64 numpy oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import grg_oracle as O
from oracle import lmm_oracle as L
from oracle import mlp_oracle as N
from tests._builders import build_hip_layout, write_synthetic_artifact

pytestmark = pytest.mark.gpu


def _req(k):
    from pygrgl_spmv.grg import RuntimeRequirements

    return RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=False, need_up_miss_output=False,
                               need_init_vector=False, need_init_matrix=False, need_init_xtx=True)


def _cfg(seed, samples, mutations):
    from pygrgl_spmv.synth import SynthConfig

    return SynthConfig(num_samples=samples, num_mutations=mutations, seed=seed, ploidy=2, num_trees=80,
                       keep_fraction=0.35, window_fraction=0.1, num_segments=2)


@pytest.fixture(scope="module")
def lmm_arts(artifact_dir):
    return [write_synthetic_artifact(artifact_dir, "lmm-a", _cfg(91, 160, 1200)),
            write_synthetic_artifact(artifact_dir, "lmm-b", _cfg(92, 160, 700))]


@pytest.fixture(scope="module")
def dense_k(lmm_arts):
    Gs = [L.dosages(O.load_artifact(p)) for p in lmm_arts]
    return L.dense_kinship(Gs)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)], ids=["fp64", "fp32c"])
def test_operator_matches_dense_kinship(lmm_arts, dense_k, dtype, tol):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.lmm import StandardizedGrgOperator

    K, m = dense_k
    with HipRuntime(build_hip_layout(lmm_arts, dtype=dtype, requirements=_req(8))) as rt:
        op = StandardizedGrgOperator(rt, k=8)
        assert op.m_proj == m and op.n == K.shape[0]
        V = np.random.default_rng(5).standard_normal((op.n, 6))
        got = op.apply_k(torch.from_numpy(V).to(op.device)).cpu().numpy()
    exp = K @ V
    assert O.normwise_rel_err(got.T, exp.T) <= tol


def test_cg_and_eigenpairs_match_dense(lmm_arts, dense_k):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.lmm import StandardizedGrgOperator, conj_grad, shifted, top_eigenpairs

    K, _ = dense_k
    n = K.shape[0]
    with HipRuntime(build_hip_layout(lmm_arts, dtype=np.float64, requirements=_req(16))) as rt:
        op = StandardizedGrgOperator(rt, k=16)
        rng = np.random.default_rng(6)
        Y = rng.standard_normal((n, 5))
        delta = np.array([0.7, 1.0, 2.5, 5.0, 9.0])
        res = conj_grad(shifted(op, torch.from_numpy(delta).to(op.device)), torch.from_numpy(Y).to(op.device), rel_tol=1e-10,
                        max_iter=400)
        eig = top_eigenpairs(op, 4, oversample=12, iters=40)
        got_x = res.x.cpu().numpy()
        got_vals = eig.values.cpu().numpy()
    for c in range(5):
        exact = L.project(np.linalg.solve(K + delta[c] * np.eye(n), L.project(Y[:, c:c + 1])))
        np.testing.assert_allclose(got_x[:, c:c + 1], exact, rtol=1e-6, atol=1e-8)
    assert (res.column_iterations.cpu().numpy() > 0).all()
    top = np.sort(np.linalg.eigvalsh(K))[::-1][:4]
    np.testing.assert_allclose(got_vals, top, rtol=1e-6)


def test_network_kernels_match_oracle():
    import torch

    from pygrgl_spmv.lmm.mlp import DenseNet

    dev = torch.device("cuda", 0)
    net = DenseNet([21, 256, 256, 21], device=dev, seed=3)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((37, 21)).astype(np.float32)
    T = rng.standard_normal((37, 21)).astype(np.float32)
    w = rng.uniform(0.5, 2.0, 21).astype(np.float32)
    Ws = [W.cpu().numpy().astype(np.float64) for W in net.W]
    bs = [b.cpu().numpy().astype(np.float64) + 0.01 for b in net.b]
    for b, bv in zip(net.b, bs):
        b.copy_(torch.from_numpy(bv.astype(np.float32)))
    out = net.forward(torch.from_numpy(X).to(dev)).cpu().numpy()
    np.testing.assert_allclose(out, N.forward(Ws, bs, X)[-1], rtol=1e-4, atol=1e-4)
    loss, grads = net.gradients(torch.from_numpy(X).to(dev), torch.from_numpy(T).to(dev), torch.from_numpy(w).to(dev))
    e_loss, dWs, dbs = N.loss_and_grads(Ws, bs, X, T, w)
    assert float(loss.item()) == pytest.approx(e_loss, rel=1e-4)
    for li in range(3):
        gW, gb = grads[2 * li].cpu().numpy(), grads[2 * li + 1].cpu().numpy()
        assert np.abs(gW - dWs[li]).max() <= 1e-4 * max(1.0, np.abs(dWs[li]).max())
        assert np.abs(gb - dbs[li]).max() <= 1e-4 * max(1.0, np.abs(dbs[li]).max())
    # two Adam steps against the oracle
    p_ref = [Ws[0].copy(), np.zeros_like(Ws[0]), np.zeros_like(Ws[0])]
    for t in (1, 2):
        net.adam_step(grads, lr=1e-2)
        p_ref[0], p_ref[1], p_ref[2] = N.adam(p_ref[0], grads[0].cpu().numpy().astype(np.float64), p_ref[1], p_ref[2], t, lr=1e-2)
    np.testing.assert_allclose(net.W[0].cpu().numpy(), p_ref[0], rtol=1e-5, atol=1e-6)


def test_network_fits_a_smooth_map():
    import torch

    from pygrgl_spmv.lmm.mlp import DenseNet

    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(8)
    X = rng.uniform(-1, 1, (512, 3)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2], X[:, 0] ** 2 - X[:, 2]], axis=1).astype(np.float32)
    net = DenseNet([3, 64, 64, 2], device=dev, seed=1)
    hist = net.train(torch.from_numpy(X).to(dev), torch.from_numpy(T).to(dev), epochs=150, batch=64, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0]


def test_warm_start_saves_iterations(artifact_dir):
    """Train the warm start on plain-CG solves of a small GRG. The deployed start (the
    network when it beat the closed-form start on held-out validation solves, otherwise
    the closed form) reaches the same residual as the zero start in fewer iterations, with
    the same solutions, and never needs more iterations than the closed-form start."""
    from pygrgl_spmv.lmm.experiment import run

    art = write_synthetic_artifact(artifact_dir, "lmm-warm", _cfg(93, 1200, 12000))
    res = run([art], k=32, r=20, n_train=512, n_test=64, rel_tol=1e-6, max_iter=400, dtype=np.float32, epochs=150)
    t = res["test"]
    print({name: (v["mean_iterations"], v.get("iterations_saved")) for name, v in t.items()})
    assert t["zero"]["unconverged"] == 0 and t["network"]["unconverged"] == 0
    assert t["network"]["max_rel_residual"] <= 1e-6
    assert t["network"]["max_rel_diff_vs_zero_start"] <= 1e-4
    assert t["network"]["mean_iterations"] < t["zero"]["mean_iterations"]
    assert t["network"]["mean_iterations"] <= t["closed_form"]["mean_iterations"]
    v = res["validation"]
    print({"validation": v, "network_raw": t["network_raw"]["mean_iterations"]})
    assert v["network_deployed"] == (v["network_mean_iterations"] <= v["closed_form_mean_iterations"] - v["margin"])


# ---- variance components (MC scaling, secant search) and LOCO solves ----

@pytest.fixture(scope="module")
def dense_std(lmm_arts):
    Gs = [L.dosages(O.load_artifact(p)) for p in lmm_arts]
    return Gs, L.standardized(Gs)


def _phenotype(std, seed):
    rng = np.random.default_rng(seed)
    Xs = [x for x, _ in std]
    m = sum(x.shape[1] for x in Xs)
    g = sum(x @ rng.standard_normal(x.shape[1]) for x in Xs) / np.sqrt(m)
    return g * np.sqrt(0.4) + rng.standard_normal(Xs[0].shape[0]) * np.sqrt(0.6)


def test_scores_and_parts_match_dense(lmm_arts, dense_std):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.lmm import StandardizedGrgOperator

    _, std = dense_std
    with HipRuntime(build_hip_layout(lmm_arts, dtype=np.float64, requirements=_req(8))) as rt:
        op = StandardizedGrgOperator(rt, k=8)
        V = torch.from_numpy(np.random.default_rng(7).standard_normal((op.n, 5))).to(op.device)
        scores = [s.cpu().numpy() for s in op.scores(V)]
        parts = op.apply_parts(V)
        K_V = op.apply_k(V).cpu().numpy()
        summed = L.project(sum(p.cpu().numpy() for p in parts) / op.m_proj)
        assert op.m_per_grg == [int(keep.sum()) for _, keep in std]
    PV = L.project(V.cpu().numpy())
    for (X, keep), sc, part in zip(std, scores, parts):
        assert O.normwise_rel_err(sc[keep].T, (X.T @ PV).T) <= 1e-12
        assert not sc[~keep].any()
        assert O.normwise_rel_err(L.project(part.cpu().numpy()).T, (X @ (X.T @ PV)).T) <= 1e-12
    assert O.normwise_rel_err(summed.T, K_V.T) <= 1e-13


def test_mc_scaling_matches_dense(lmm_arts, dense_std):
    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.lmm import StandardizedGrgOperator
    from pygrgl_spmv.lmm import variance as V

    _, std = dense_std
    Xs, masks = [x for x, _ in std], [k for _, k in std]
    y = _phenotype(std, 11)
    trials = 6
    with HipRuntime(build_hip_layout(lmm_arts, dtype=np.float64, requirements=_req(8))) as rt:
        op = StandardizedGrgOperator(rt, k=8)
        G, E, ydev = V.mc_components(op, y, trials=trials, seed=3)
        betas, noise = V.mc_draws(op.scales, op.n, op.m_proj, trials=trials, seed=3)
        np.testing.assert_allclose(G.cpu().numpy(), L.genetic_values(Xs, masks, betas), atol=1e-11)
        np.testing.assert_allclose(E.cpu().numpy(), L.project(noise), atol=1e-13)
        for h2 in (0.25, 0.6):
            ld = V.log_delta_from_h2(op, h2)
            assert V.h2_from_log_delta(op, ld) == pytest.approx(h2, rel=1e-12)
            got = V.mc_scaling(op, ydev, G, E, log_delta=ld, rel_tol=1e-12, max_iter=1000)
            exp = L.mc_scaling(Xs, L.genetic_values(Xs, masks, betas), noise, y, ld)
            assert got.f_reml == pytest.approx(exp["f_reml"], abs=1e-9)
            assert got.f_jacks == pytest.approx(tuple(exp["f_jacks"]), abs=1e-9)
            assert got.f_rands_as_data == pytest.approx(tuple(exp["f_rands_as_data"]), abs=1e-9)
            assert got.sigma2_k == pytest.approx(exp["sigma2_k"], rel=1e-9)
            np.testing.assert_allclose(got.hinv_y.cpu().numpy(), exp["hinv_y"], rtol=0, atol=1e-9 * np.abs(exp["hinv_y"]).max())


def test_variance_fit_and_loco_match_dense(lmm_arts, dense_std):
    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.lmm import StandardizedGrgOperator
    from pygrgl_spmv.lmm import variance as V

    _, std = dense_std
    Xs, masks = [x for x, _ in std], [k for _, k in std]
    y = _phenotype(std, 12)
    with HipRuntime(build_hip_layout(lmm_arts, dtype=np.float64, requirements=_req(8))) as rt:
        op = StandardizedGrgOperator(rt, k=8)
        fit = V.fit_variance_components(op, y, mc_trials=5, seed=4, rel_tol=1e-12, max_iter=1000)
        loco = V.solve_loco(op, y, delta=fit.delta, rel_tol=1e-12, max_iter=1000).cpu().numpy()
        betas, noise = V.mc_draws(op.scales, op.n, op.m_proj, trials=5, seed=4)
    G = L.genetic_values(Xs, masks, betas)
    assert 2 <= len(fit.evaluations) <= 7
    for ev in fit.evaluations:                                   # every point of the search against the dense solve
        assert ev.f_reml == pytest.approx(L.mc_scaling(Xs, G, noise, y, ev.log_delta)["f_reml"], abs=1e-9)
    secant = fit.evaluations[2:] or fit.evaluations
    assert fit.log_delta in [e.log_delta for e in fit.evaluations]
    assert abs(next(e for e in fit.evaluations if e.log_delta == fit.log_delta).f_reml) <= min(abs(e.f_reml) for e in secant) + 1e-15
    assert fit.h2 == pytest.approx(V.h2_from_log_delta(op, fit.log_delta)) and fit.sigma_e2 == pytest.approx(fit.delta * fit.sigma_g2)
    exp = L.loco_solutions(Xs, y, fit.delta)
    np.testing.assert_allclose(loco, exp, rtol=0, atol=1e-9 * np.abs(exp).max())
