"""CPU checks of the mixed-model side: the network oracle, the dense LMM oracle,
the MLP library's exports (no GPU calls)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import grg_oracle as O
from oracle import lmm_oracle as L
from oracle import mlp_oracle as N


def _net(rng, sizes):
    Ws = [rng.standard_normal((o, i)) * 0.5 for i, o in zip(sizes[:-1], sizes[1:])]
    bs = [rng.standard_normal(o) * 0.1 for o in sizes[1:]]
    return Ws, bs


def test_network_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    Ws, bs = _net(rng, [5, 7, 6, 3])
    X = rng.standard_normal((9, 5))
    T = rng.standard_normal((9, 3))
    w = np.array([1.0, 0.5, 2.0])
    _, dWs, dbs = N.loss_and_grads(Ws, bs, X, T, w)
    h = 1e-6
    for li in range(3):
        for idx in [(0, 0), (1, 2), (Ws[li].shape[0] - 1, Ws[li].shape[1] - 1)]:
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[li][idx] += h
            Wm[li][idx] -= h
            fd = (N.loss_and_grads(Wp, bs, X, T, w)[0] - N.loss_and_grads(Wm, bs, X, T, w)[0]) / (2 * h)
            assert dWs[li][idx] == pytest.approx(fd, rel=1e-6, abs=1e-8)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[li][0] += h
        bm[li][0] -= h
        fd = (N.loss_and_grads(Ws, bp, X, T, w)[0] - N.loss_and_grads(Ws, bm, X, T, w)[0]) / (2 * h)
        assert dbs[li][0] == pytest.approx(fd, rel=1e-6, abs=1e-8)


def test_adam_oracle_first_step_is_signed_lr():
    """With zero moments, step 1 moves every parameter by lr * g / (|g| + eps) ~ lr * sign(g)."""
    g = np.array([0.3, -2.0, 5e-3])
    p, m, v = N.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(p, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


@pytest.fixture(scope="module")
def tiny_grg(tmp_path_factory):
    from pygrgl_spmv.synth import SynthConfig, write_artifact

    cfg = SynthConfig(num_samples=120, num_mutations=800, seed=77, ploidy=2, num_trees=60, keep_fraction=0.35,
                      window_fraction=0.1, num_segments=2)
    path = write_artifact(cfg, tmp_path_factory.mktemp("lmm") / "tiny.grg_spmv")
    return O.load_artifact(path)


def test_dosages_match_dense_genotypes(tiny_grg):
    X = O.dense_genotypes(tiny_grg)
    G = np.zeros((tiny_grg.num_individuals, tiny_grg.num_mutations))
    np.add.at(G, tiny_grg.sample_to_individual, X)
    np.testing.assert_array_equal(L.dosages(tiny_grg), G)


def test_kinship_oracle_invariants(tiny_grg):
    G = L.dosages(tiny_grg)
    K, m = L.dense_kinship([G])
    n = G.shape[0]
    assert m == int((L.variant_stats(G)[2] > 0).sum())
    np.testing.assert_allclose(K, K.T, atol=1e-12)
    assert np.trace(K) == pytest.approx(n - 1.0, rel=1e-12)       # every standardized column has norm^2 n - 1
    np.testing.assert_allclose(K @ np.ones(n), 0.0, atol=1e-9)      # the covariate (all-ones) direction is projected out
    assert np.linalg.eigvalsh(K).min() > -1e-9


def test_cg_oracle_matches_direct_solve(tiny_grg):
    G = L.dosages(tiny_grg)
    K, _ = L.dense_kinship([G])
    n = G.shape[0]
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal((n, 4))
    delta = 0.8
    H = K + delta * np.eye(n)
    x, it = L.conj_grad(H, rhs, rel_tol=1e-10, max_iter=500)
    exact = L.project(np.linalg.solve(H, L.project(rhs)))
    np.testing.assert_allclose(x, exact, rtol=1e-8, atol=1e-8)
    assert 1 <= it <= n


def test_mlp_library_exports():
    from pygrgl_spmv.lmm import mlp

    if not mlp.LIB_PATH.exists():
        pytest.skip("libgrg_mlp.so not built")
    lib = ctypes.CDLL(str(mlp.LIB_PATH))
    for name in ("grg_mlp_linear_forward", "grg_mlp_linear_backward", "grg_mlp_mse", "grg_mlp_adam",
                 "grg_mlp_last_error", "grg_mlp_abi_version"):
        assert hasattr(lib, name), name
    lib.grg_mlp_abi_version.restype = ctypes.c_int32
    assert lib.grg_mlp_abi_version() == mlp.ABI_VERSION
    m = mlp.MlpLibrary()
    with pytest.raises(RuntimeError, match="HIP grg_mlp_adam failed: status 1: Adam step"):
        m.check(m.lib.grg_mlp_adam(ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 1,
                                   0.1, 0.9, 0.999, 1e-8, 0, None), "grg_mlp_adam")


# ---- variance components and LOCO (oracle pins and host logic) ----

@pytest.fixture(scope="module")
def two_grgs(tmp_path_factory):
    from pygrgl_spmv.synth import SynthConfig, write_artifact

    d = tmp_path_factory.mktemp("vc")
    out = []
    for seed, m in ((81, 600), (82, 400)):
        cfg = SynthConfig(num_samples=100, num_mutations=m, seed=seed, ploidy=2, num_trees=50, keep_fraction=0.35,
                          window_fraction=0.1, num_segments=2)
        out.append(L.dosages(O.load_artifact(write_artifact(cfg, d / f"g{seed}.grg_spmv"))))
    return out


def test_score_square_sums_equal_m_ztkz(two_grgs):
    std = L.standardized(two_grgs)
    Xs = [x for x, _ in std]
    K = L.kinship_of(Xs)
    m = sum(x.shape[1] for x in Xs)
    Z = np.random.default_rng(1).standard_normal((K.shape[0], 3))
    PZ = L.project(Z)
    np.testing.assert_allclose(L.score_square_sums(Xs, Z), m * (PZ * (K @ PZ)).sum(axis=0), rtol=1e-10)
    np.testing.assert_allclose(K, L.dense_kinship(two_grgs)[0], atol=1e-12)


def test_mc_scaling_oracle_is_zero_for_isotropic_kinship():
    """K = c P (orthonormal centred columns): beta / eps = M c for every column, so f_REML = 0."""
    rng = np.random.default_rng(2)
    n, m = 40, 39
    Q, _ = np.linalg.qr(L.project(rng.standard_normal((n, m))))
    Xs = [Q[:, :20] * 3.0, Q[:, 20:] * 3.0]
    G, E, y = rng.standard_normal((n, 4)), rng.standard_normal((n, 4)), rng.standard_normal(n)
    for ld in (-1.0, 0.3, 2.0):
        r = L.mc_scaling(Xs, G, E, y, ld)
        assert abs(r["f_reml"]) < 1e-10
        np.testing.assert_allclose(r["f_jacks"], 0.0, atol=1e-10)


def test_mc_scaling_oracle_data_equal_to_a_trial(two_grgs):
    Xs = [x for x, _ in L.standardized(two_grgs)]
    n = Xs[0].shape[0]
    rng = np.random.default_rng(3)
    G, E = rng.standard_normal((n, 5)), rng.standard_normal((n, 5))
    ld = 0.4
    y = E[:, 2] * np.sqrt(np.exp(ld)) + G[:, 2]
    r = L.mc_scaling(Xs, G, E, y, ld)
    assert r["f_reml"] == pytest.approx(r["f_rands_as_data"][2], abs=1e-10)
    H = L.kinship_of(Xs) + np.exp(ld) * np.eye(n)
    np.testing.assert_allclose(H @ r["hinv_y"], L.project(y.reshape(n, 1))[:, 0], atol=1e-9)


def test_loco_oracle_two_grgs_is_the_other_kinship(two_grgs):
    std = L.standardized(two_grgs)
    Xs = [x for x, _ in std]
    n = Xs[0].shape[0]
    y = np.random.default_rng(4).standard_normal(n)
    Z = L.loco_solutions(Xs, y, 0.9)
    for c in range(2):
        K_other, _ = L.dense_kinship([two_grgs[1 - c]])
        np.testing.assert_allclose(Z[:, c], L.project(np.linalg.solve(K_other + 0.9 * np.eye(n), L.project(y[:, None])))[:, 0],
                                   atol=1e-10)


def test_reml_objective_known_values():
    from pygrgl_spmv.lmm.variance import reml_objective

    f, jacks, as_data = reml_objective(2.0, 1.0, [1.0, 3.0], [1.0, 3.0])
    assert f == pytest.approx(np.log(2.0))
    assert jacks == pytest.approx((np.log(2.0), np.log(2.0), np.log(2.0)))
    assert as_data == pytest.approx((0.0, 0.0))
    f, jacks, _ = reml_objective(1.0, 1.0, [1.0, 4.0], [2.0, 2.0])   # sum beta / sum eps = 5/4
    assert f == pytest.approx(np.log(0.8))
    assert jacks[0] == pytest.approx(np.log(1.0 / 2.0)) and jacks[1] == pytest.approx(np.log(2.0))
    with pytest.raises(RuntimeError, match="invalid MC-scaling"):
        reml_objective(1.0, 0.0, [1.0], [1.0])


class _Eval:
    def __init__(self, f):
        self.f, self.calls = f, []

    def __call__(self, ld):
        from types import SimpleNamespace

        self.calls.append(ld)
        return SimpleNamespace(log_delta=ld, f_reml=self.f(ld))


def test_secant_search_finds_a_linear_root_and_clips():
    from pygrgl_spmv.lmm.variance import secant_search

    at = {0.25: 1.0, 0.125: 2.0, 0.5: 0.0}.__getitem__       # log delta falls as h2 rises
    ev = _Eval(lambda ld: 0.7 * (1.6 - ld))                   # f(1) > 0 -> second start at h2 = 0.5
    best, tried = secant_search(ev, at)
    assert ev.calls[:2] == [1.0, 0.0] and best.log_delta == pytest.approx(1.6, abs=1e-12)
    assert 3 <= len(tried) <= 7                                # two starts, at most five secant steps
    ev = _Eval(lambda ld: 0.1 - 0.01 * ld)                     # root at 10: start at h2 = 0.5 then secant
    best, _ = secant_search(ev, at)
    assert best.log_delta == pytest.approx(10.0)
    ev = _Eval(lambda ld: 0.1 - 0.001 * ld)                    # root at 100: clipped to 10
    best, _ = secant_search(ev, at)
    assert max(ev.calls) == 10.0 and best.log_delta == 10.0
    ev = _Eval(lambda ld: ld - 1.3)                            # f(1) < 0 -> second start at h2 = 0.125
    best, _ = secant_search(ev, at)
    assert ev.calls[1] == 2.0 and best.log_delta == pytest.approx(1.3)


def test_mc_draws_zero_rows_and_shapes():
    from pygrgl_spmv.lmm.variance import mc_draws

    scales = [np.array([1.0, 0.0, 2.0]), np.array([0.0, 0.5])]
    betas, noise = mc_draws(scales, 7, 3, trials=4, seed=5)
    assert [b.shape for b in betas] == [(3, 4), (2, 4)] and noise.shape == (7, 4)
    assert not betas[0][1].any() and not betas[1][0].any() and betas[0][0].all()
    _, n2 = mc_draws(scales, 7, 3, trials=4, seed=5)
    np.testing.assert_array_equal(n2, noise)
