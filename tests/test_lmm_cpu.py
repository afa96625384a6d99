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
