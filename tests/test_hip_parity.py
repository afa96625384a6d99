"""HIP sweep parity against the CPU oracle (GPU).

Bar: FP64 results match the oracle within 1e-12 normwise relative error
(per output column, max|a - b| / max|b|) and exactly for integer-valued
inputs (+-1, all-ones, 0/1: sums of integers below 2^53 are order-independent);
FP32 (compensated) results within 5e-7 normwise relative error (the FP32 storage
floor is 2^-24 per stored node value; measured errors are 1e-7 to 3e-7).
Mode matrix follows the reference's ``runtime/test_prepare_cuda.py:124-148``
(direction x by_individual x init mode), plus missing data, emit_all_nodes,
a k sweep like ``runtime/test_traversal_correctness.py:14,59-74`` and the band
closed forms of ``test_triton.py:852-895``.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import grg_oracle as O
from tests._builders import build_hip_layout, full_requirements, write_band_artifact, write_synthetic_artifact

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 5e-7


def _cfg_small():
    from pygrgl_spmv.synth import SynthConfig

    return SynthConfig(num_samples=400, num_mutations=3000, seed=101, ploidy=2, num_trees=150, keep_fraction=0.3,
                       window_fraction=0.1, num_segments=3, missing_fraction=0.05, missing_pool=8)


@pytest.fixture(scope="module")
def small_artifact(artifact_dir):
    return write_synthetic_artifact(artifact_dir, "parity-small", _cfg_small())


@pytest.fixture(scope="module")
def small_oracle(small_artifact):
    return O.load_artifact(small_artifact)


@pytest.fixture(scope="module")
def runtime64(small_artifact):
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([small_artifact], requirements=full_requirements(max_k_up=16, max_k_down=16))
    with HipRuntime(layout) as rt:
        yield rt


def _cols(g, direction, by_individual):
    if direction == "up":
        return g.num_individuals if by_individual else g.num_samples
    return g.num_mutations


def _init(rng, g, mode, k):
    if mode == "none":
        return None
    if mode == "vector":
        return rng.standard_normal(k)
    if mode == "matrix":
        return rng.standard_normal((k, g.num_nodes))
    return "xtx"


@pytest.mark.parametrize("direction", ["up", "down"])
@pytest.mark.parametrize("by_individual", [False, True], ids=["by-node", "by-individual"])
@pytest.mark.parametrize("init_mode", ["none", "vector", "matrix", "xtx"])
def test_modes_match_oracle(runtime64, small_oracle, direction, by_individual, init_mode):
    (grg,) = runtime64.grgs
    g = small_oracle
    k = 2
    rng = np.random.default_rng(72_100 + (100 if direction == "down" else 0) + (10 if by_individual else 0) + len(init_mode))
    x = rng.standard_normal((k, _cols(g, direction, by_individual)))
    init = _init(rng, g, init_mode, k)
    expected = O.matmul(g, x, direction, by_individual=by_individual, init=init)
    actual = grg.matmul(x, direction, by_individual=by_individual, init=init)
    assert actual.shape == expected.shape
    assert O.normwise_rel_err(actual, expected) <= FP64_TOL


@pytest.mark.parametrize("k", [1, 2, 3, 7, 8, 9, 16])
@pytest.mark.parametrize("direction", ["up", "down"])
def test_k_sweep_exact_pm1(runtime64, small_oracle, k, direction):
    (grg,) = runtime64.grgs
    rng = np.random.default_rng(8000 + k)
    x = rng.choice(np.array([-1.0, 1.0]), size=(k, _cols(small_oracle, direction, False)))
    np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(small_oracle, x, direction))


def test_all_ones_allele_counts(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    ones = np.ones((1, small_oracle.num_samples))
    counts = grg.matmul(ones, "up")
    np.testing.assert_array_equal(counts, O.matmul(small_oracle, ones, "up"))
    X = O.dense_genotypes(small_oracle)
    np.testing.assert_array_equal(counts[0], X.sum(axis=0))


def test_xtx_closed_form(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    got = grg.matmul(np.ones((2, small_oracle.num_samples)), "up", init="xtx")
    np.testing.assert_array_equal(got[0], O.xtx_closed_form(small_oracle))
    np.testing.assert_array_equal(got[1], got[0])


@pytest.mark.parametrize("direction", ["up", "down"])
def test_emit_all_nodes(runtime64, small_oracle, direction):
    (grg,) = runtime64.grgs
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, _cols(small_oracle, direction, False)))
    got = grg.matmul(x, direction, emit_all_nodes=True)
    exp = O.matmul(small_oracle, x, direction, emit_all_nodes=True)
    assert got.shape == (3, small_oracle.num_nodes)
    assert O.normwise_rel_err(got, exp) <= FP64_TOL


def test_emit_all_nodes_by_individual_up(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    x = np.random.default_rng(6).standard_normal((2, small_oracle.num_individuals))
    got = grg.matmul(x, "up", emit_all_nodes=True, by_individual=True)
    exp = O.matmul(small_oracle, x, "up", emit_all_nodes=True, by_individual=True)
    assert O.normwise_rel_err(got, exp) <= FP64_TOL


def test_missing_up_output(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    rng = np.random.default_rng(72_400)
    x = rng.standard_normal((2, small_oracle.num_individuals))
    miss_got = rng.standard_normal((2, small_oracle.num_mutations))
    miss_exp = miss_got.copy()
    got = grg.matmul(x, "up", by_individual=True, miss=miss_got)
    exp = O.matmul(small_oracle, x, "up", by_individual=True, miss=miss_exp)
    assert O.normwise_rel_err(got, exp) <= FP64_TOL
    assert O.normwise_rel_err(miss_got, miss_exp) <= FP64_TOL


def test_missing_down_input(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    rng = np.random.default_rng(72_401)
    y = rng.standard_normal((3, small_oracle.num_mutations))
    miss = rng.standard_normal((3, small_oracle.num_mutations))
    got = grg.matmul(y, "down", miss=miss)
    exp = O.matmul(small_oracle, y, "down", miss=miss)
    assert O.normwise_rel_err(got, exp) <= FP64_TOL


def test_zero_input_gives_zero(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    np.testing.assert_array_equal(grg.matmul(np.zeros((4, small_oracle.num_samples)), "up"), 0.0)
    np.testing.assert_array_equal(grg.matmul(np.zeros((4, small_oracle.num_mutations)), "down"), 0.0)


def test_repeat_stability(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    x = np.random.default_rng(11).standard_normal((5, small_oracle.num_samples))
    first = grg.matmul(x, "up")
    for _ in range(3):
        np.testing.assert_array_equal(grg.matmul(x, "up"), first)


def test_adjointness(runtime64, small_oracle):
    (grg,) = runtime64.grgs
    rng = np.random.default_rng(12)
    x = rng.standard_normal((1, small_oracle.num_samples))
    y = rng.standard_normal((1, small_oracle.num_mutations))
    lhs = float((y * grg.matmul(x, "up")).sum())
    rhs = float((grg.matmul(y, "down") * x).sum())
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)


def test_fp32_compensated_matches_oracle(small_artifact, small_oracle):
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([small_artifact], dtype=np.float32, requirements=full_requirements(max_k_up=16, max_k_down=16))
    rng = np.random.default_rng(31)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        for direction in ("up", "down"):
            x = rng.standard_normal((16, _cols(small_oracle, direction, False))).astype(np.float32)
            got = grg.matmul(x, direction)
            assert got.dtype == np.float32
            exp = O.matmul(small_oracle, x.astype(np.float64), direction)
            assert O.normwise_rel_err(got, exp) <= FP32_TOL
        ones = np.ones((4, small_oracle.num_samples), dtype=np.float32)
        np.testing.assert_array_equal(grg.matmul(ones, "up"), O.matmul(small_oracle, ones.astype(np.float64), "up"))


@pytest.mark.parametrize("order", [("up", "down"), ("down", "up")], ids=["up-down", "down-up"])
@pytest.mark.parametrize("k", [1, 2])
def test_band_closed_form_exact(artifact_dir, order, k):
    from pygrgl_spmv.backends.hip import HipRuntime

    n, bw = 64, 8
    art = write_band_artifact(artifact_dir, "band-n64", n=n, bandwidth=bw)
    layout = build_hip_layout([art], requirements=full_requirements(max_k_up=k, max_k_down=k))
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        for i, direction in enumerate(order):
            rng = np.random.default_rng(61_000 + 1000 * k + 100 * i)
            x = rng.choice(np.array([-1.0, 1.0]), size=(k, n))
            if direction == "up":
                exp = O.band_expected_up(x.T, shifts=(0, bw, 2 * bw), bandwidth=bw).T
            else:
                exp = O.band_expected_down(x.T, shifts=(0, bw, 2 * bw), bandwidth=bw).T
            np.testing.assert_array_equal(grg.matmul(x, direction), exp)
