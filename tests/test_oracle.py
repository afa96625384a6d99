"""Pins the CPU oracle (no GPU).

Each check is independent of the oracle's sparse sweep arithmetic:

* band closed forms -- prefix-sum window sums, the known answers of the
  reference's exactness tests (``_streaming_cases.py:493-526``, used with exact
  equality at ``test_triton.py:852-895``);
* dense brute force -- path counts from the full adjacency by fixed-point
  iteration, no level order (``endtoend/conftest.py:38-70`` builds the dense
  genotype matrix the same way from the DAG);
* the diploid X^T X closed form for ``init='xtx'``
  (``endtoend/test_matmul.py:96-117``);
* identities: adjointness, linearity in the init vector, zero in -> zero out,
  matrix-init locality (``endtoend/test_matmul.py:120-166``);
* MKL's sparse BLAS on the same blocks (the reference's MklRuntime arithmetic).

A planted mistake (transposed block, dropped selector, wrong sweep order) makes
one of these fail -- see ``test_planted_mistakes_are_caught``.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import grg_oracle as O
from tests._builders import write_band_artifact


def _tiny_cfg(seed=11, ploidy=2, missing=0.05):
    from pygrgl_spmv.synth import SynthConfig

    return SynthConfig(num_samples=120, num_mutations=700, seed=seed, ploidy=ploidy, num_trees=60, keep_fraction=0.3,
                       window_fraction=0.15, num_segments=2, missing_fraction=missing, missing_pool=4)


@pytest.fixture(scope="module")
def tiny():
    from pygrgl_spmv.synth import generate_state

    return O.from_state(generate_state(_tiny_cfg()))


@pytest.mark.parametrize("n,bw", [(4, 1), (16, 4), (64, 8)])
@pytest.mark.parametrize("k", [1, 3])
def test_band_closed_forms(tmp_path, n, bw, k):
    art = write_band_artifact(tmp_path, f"band-{n}-{bw}", n=n, bandwidth=bw)
    g = O.load_artifact(art)
    rng = np.random.default_rng(n * 10 + k)
    x = rng.choice(np.array([-1.0, 1.0]), size=(k, n))
    shifts = (0, bw, 2 * bw)
    np.testing.assert_array_equal(O.matmul(g, x, "up"), O.band_expected_up(x.T, shifts=shifts, bandwidth=bw).T)
    np.testing.assert_array_equal(O.matmul(g, x, "down"), O.band_expected_down(x.T, shifts=shifts, bandwidth=bw).T)


def test_dense_brute_force(tiny):
    X = O.dense_genotypes(tiny)
    assert set(np.unique(X)) <= {0.0, 1.0}
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, tiny.num_samples))
    y = rng.standard_normal((4, tiny.num_mutations))
    np.testing.assert_allclose(O.matmul(tiny, x, "up"), x @ X, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.matmul(tiny, y, "down"), y @ X.T, rtol=1e-12, atol=1e-12)
    # diploid by_individual (endtoend/test_matmul.py:72-91)
    Xd = np.zeros((tiny.num_individuals, tiny.num_mutations))
    np.add.at(Xd, tiny.sample_to_individual, X)
    xi = rng.standard_normal((3, tiny.num_individuals))
    np.testing.assert_allclose(O.matmul(tiny, xi, "up", by_individual=True), xi @ Xd, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.matmul(tiny, y, "down", by_individual=True), y @ Xd.T, rtol=1e-12, atol=1e-12)


def test_xtx_closed_form(tiny):
    got = O.matmul(tiny, np.ones((2, tiny.num_samples)), "up", init="xtx")
    np.testing.assert_array_equal(got[0], O.xtx_closed_form(tiny))
    np.testing.assert_array_equal(got[1], got[0])


def test_missing_selector_brute_force(tiny):
    P = O.dense_path_counts(tiny)
    Xmiss = (tiny.sel_miss.toarray() @ P[:, : tiny.num_samples]).T
    rng = np.random.default_rng(4)
    x = rng.standard_normal((2, tiny.num_samples))
    miss = np.zeros((2, tiny.num_mutations))
    O.matmul(tiny, x, "up", miss=miss)
    np.testing.assert_allclose(miss, x @ Xmiss, rtol=1e-12, atol=1e-12)
    y = rng.standard_normal((2, tiny.num_mutations))
    m = rng.standard_normal((2, tiny.num_mutations))
    X = O.dense_genotypes(tiny)
    np.testing.assert_allclose(O.matmul(tiny, y, "down", miss=m), y @ X.T + m @ Xmiss.T, rtol=1e-12, atol=1e-12)


def test_identities(tiny):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1, tiny.num_samples))
    y = rng.standard_normal((1, tiny.num_mutations))
    assert abs(float((y * O.matmul(tiny, x, "up")).sum()) - float((O.matmul(tiny, y, "down") * x).sum())) < 1e-10
    c = np.array([2.5, -1.0])
    xx = rng.standard_normal((2, tiny.num_samples))
    vu, vd, _, _ = O.init_biases(tiny)
    np.testing.assert_allclose(O.matmul(tiny, xx, "up", init=c), O.matmul(tiny, xx, "up") + c[:, None] * vu[None, :], rtol=1e-12, atol=1e-12)
    yy = rng.standard_normal((2, tiny.num_mutations))
    np.testing.assert_allclose(O.matmul(tiny, yy, "down", init=c), O.matmul(tiny, yy, "down") + c[:, None] * vd[None, :], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(O.matmul(tiny, np.zeros((3, tiny.num_samples)), "up"), 0.0)
    # init vector monotonicity on all-ones (endtoend/test_matmul.py:120-131)
    ones = np.ones((3, tiny.num_samples))
    assert np.all(O.matmul(tiny, ones, "up", init=np.full(3, 2.0)) >= 2 * O.matmul(tiny, ones, "up"))


def test_matrix_init_locality(tiny):
    """A bump at one node reaches exactly the mutations whose nodes are at or above it."""
    ones = np.ones((2, tiny.num_samples))
    init = np.zeros((2, tiny.num_nodes))
    base = O.matmul(tiny, ones, "up", init=init)
    np.testing.assert_array_equal(base, O.matmul(tiny, ones, "up"))
    node_orig = int(tiny.node_perm[tiny.num_samples + 3])  # an internal node (original id)
    init[1, node_orig] = 100.0
    bumped = O.matmul(tiny, ones, "up", init=init)
    P = O.dense_path_counts(tiny)  # internal ids
    internal = int(np.flatnonzero(tiny.node_perm == node_orig)[0])
    above = (tiny.sel_mut.toarray() @ P[:, internal]) > 0
    np.testing.assert_array_equal(bumped[0], base[0])
    assert np.all(bumped[1][above] > base[1][above])
    np.testing.assert_array_equal(bumped[1][~above], base[1][~above])


def test_stored_biases_match_oracle(tmp_path):
    from pygrgl_spmv.synth import write_artifact

    art = write_artifact(_tiny_cfg(seed=12), tmp_path / "t.grg_spmv")
    g = O.load_artifact(art)
    vu, vd, xu, xd = O.init_biases(g)
    np.testing.assert_allclose(g.vec_up_bias, vu, rtol=1e-12)
    np.testing.assert_allclose(g.vec_down_bias, vd, rtol=1e-12)
    np.testing.assert_allclose(g.xtx_up_bias, xu, rtol=1e-12)
    np.testing.assert_allclose(g.xtx_down_bias, xd, rtol=1e-12)


def test_mkl_sweep_matches_oracle(tiny):
    from oracle.mkl_baseline import MklSweep

    try:
        sweep = MklSweep(tiny, max_k=4, n_threads=2)
    except RuntimeError as exc:
        pytest.skip(str(exc))
    try:
        rng = np.random.default_rng(8)
        for k in (1, 4):
            x = rng.standard_normal((tiny.num_samples, k))
            v = np.zeros((tiny.num_nodes, k))
            v[: tiny.num_samples] = x
            sweep.sweep(v, "up")
            assert O.normwise_rel_err(v.T, O.sweep_up(tiny, x).T) <= 1e-14
            y = rng.standard_normal((tiny.num_mutations, k))
            v = np.ascontiguousarray(tiny.sel_mut.T @ y)
            sweep.sweep(v, "down")
            assert O.normwise_rel_err(v.T, O.sweep_down(tiny, y).T) <= 1e-14
    finally:
        sweep.close()


def test_planted_mistakes_are_caught(tmp_path):
    """Mutate the oracle's input in plausible ways; the closed form must notice each one."""
    n, bw = 16, 4
    art = write_band_artifact(tmp_path, "band-plant", n=n, bandwidth=bw)
    g = O.load_artifact(art)
    x = np.random.default_rng(9).choice(np.array([-1.0, 1.0]), size=(2, n))
    exp_up = O.band_expected_up(x.T, shifts=(0, bw, 2 * bw), bandwidth=bw).T
    exp_down = O.band_expected_down(x.T, shifts=(0, bw, 2 * bw), bandwidth=bw).T
    np.testing.assert_array_equal(O.matmul(g, x, "up"), exp_up)
    # transposed block
    g.blocks[(2, 1)] = g.blocks[(2, 1)].T.tocsr()
    assert not np.array_equal(O.matmul(g, x, "up"), exp_up)
    g = O.load_artifact(art)
    # dropped block
    del g.blocks[(2, 0)]
    assert not np.array_equal(O.matmul(g, x, "down"), exp_down)
    g = O.load_artifact(art)
    # wrong shift in the closed form itself
    assert not np.array_equal(O.band_expected_up(x.T, shifts=(0, bw, 2 * bw + 1), bandwidth=bw).T, exp_up)
