"""The HIP sweeps on GRGs read from the reference's own test data
(``tests/golden/grgl``, converted by ``pygrgl_spmv.convert``), against dense
genotype matrices: the 200-sample GRG against its IGD genotype file, the
msprime example (the reference's primary test GRG) against the matrix of its
own child lists. Mirrors the reference's ``tests/endtoend/test_matmul.py``
(methods, diploid, xtx / vector / matrix init, init failures) plus the
missing-data input and output. FP64, exact inputs or 1e-12 relative."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import grg_oracle as O
from pygrgl_spmv import convert
from pygrgl_spmv.grg.grgl_file import read_grg
from tests._builders import build_hip_layout, full_requirements
from tests._igd import genotype_columns, read_igd

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "grgl"
T200 = GOLDEN / "test-200-samples.miss.final.grg"
MSPRIME = GOLDEN / "msprime.example.igd.final.grg"
IGD = GOLDEN / "test-200-samples.miss.igd"


def _samples_below(g):
    below = []
    for v in range(g.num_nodes):
        if v < g.num_samples:
            below.append(np.array([v]))
        else:
            kids = g.child_idx[g.child_ptr[v]:g.child_ptr[v + 1]]
            below.append(np.concatenate([below[c] for c in kids]) if kids.size else np.zeros(0, np.int64))
    return below


@pytest.fixture(scope="module")
def fixtures(tmp_path_factory):
    d = tmp_path_factory.mktemp("grgl-gpu")
    t200 = read_grg(T200)
    X, Miss = genotype_columns(read_igd(IGD), t200.positions, t200.ref_alleles, t200.alleles)
    ms = read_grg(MSPRIME)
    below = _samples_below(ms)
    Xm = np.zeros((ms.num_samples, ms.num_mutations))
    for j, n in enumerate(ms.mutation_node):
        if n >= 0:
            np.add.at(Xm[:, j], below[n], 1.0)      # path counts: a sample reached twice counts twice
    return {
        "t200": (convert(T200, d, name="t200"), t200, X, Miss),
        "msprime": (convert(MSPRIME, d, name="msprime"), ms, Xm, np.zeros_like(Xm)),
    }


def _diploid(X, ploidy=2):
    D = np.zeros((X.shape[0] // ploidy, X.shape[1]))
    np.add.at(D, np.arange(X.shape[0]) // ploidy, X)
    return D


@pytest.fixture(scope="module", params=["t200", "msprime"])
def case(request, fixtures):
    from pygrgl_spmv.backends.hip import HipRuntime

    path, g, X, Miss = fixtures[request.param]
    rt = HipRuntime(build_hip_layout([path], dtype=np.float64, requirements=full_requirements(max_k_up=8, max_k_down=8)))
    rt.__enter__()
    yield rt.grgs[0], g, X, Miss
    rt.__exit__(None, None, None)


def test_different_methods(case):
    grg, g, X, _ = case
    rng = np.random.default_rng(42)
    up_in = rng.random((8, g.num_samples))
    np.testing.assert_allclose(grg.matmul(up_in, "up"), up_in @ X, rtol=1e-12, atol=1e-12)
    down_in = rng.random((8, g.num_mutations))
    np.testing.assert_allclose(grg.matmul(down_in, "down"), down_in @ X.T, rtol=1e-12, atol=1e-12)


def test_diploid(case):
    grg, g, X, _ = case
    D = _diploid(X)
    rng = np.random.default_rng(100)
    a = rng.random((6, g.num_individuals))
    np.testing.assert_allclose(grg.matmul(a, "up", by_individual=True), a @ D, rtol=1e-12, atol=1e-12)
    b = rng.random((6, g.num_mutations))
    np.testing.assert_allclose(grg.matmul(b, "down", by_individual=True), b @ D.T, rtol=1e-12, atol=1e-12)


def test_exact_counts_and_xtx_init(case):
    grg, g, X, _ = case
    ones = np.ones((2, g.num_samples))
    np.testing.assert_array_equal(grg.matmul(ones, "up"), np.tile(X.sum(axis=0), (2, 1)))
    D = _diploid(X)
    np.testing.assert_array_equal(grg.matmul(ones, "up", init="xtx"), np.tile((D * D).sum(axis=0), (2, 1)))


def test_vector_and_matrix_init(case):
    grg, g, X, _ = case
    rows = 8
    ones = np.ones((rows, g.num_samples))
    without = grg.matmul(ones, "up")
    with_vec = grg.matmul(ones, "up", init=np.full(rows, 2.0))
    assert np.all(with_vec >= 2 * without)
    zero_init = np.zeros((rows, g.num_nodes))
    np.testing.assert_array_equal(grg.matmul(ones, "up", init=zero_init), without)


def test_missing_data_output_and_input(case):
    grg, g, X, Miss = case
    rng = np.random.default_rng(7)
    a = rng.random((4, g.num_samples))
    miss = np.zeros((4, g.num_mutations))
    out = grg.matmul(a, "up", miss=miss)
    np.testing.assert_allclose(out, a @ X, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(miss, a @ Miss, rtol=1e-12, atol=1e-12)
    y = rng.random((4, g.num_mutations))
    m = rng.random((4, g.num_mutations))
    np.testing.assert_allclose(grg.matmul(y, "down", miss=m), y @ X.T + m @ Miss.T, rtol=1e-12, atol=1e-12)


def test_init_failures(case):
    grg, g, _, _ = case
    bad = [
        dict(x=np.ones((1, g.num_samples)), init=np.ones((1, 200000))),
        dict(x=np.ones((1, g.num_samples)), init=np.ones((2, g.num_samples))),
        dict(x=np.ones((1, g.num_samples), dtype=np.int32), init=np.ones((1, g.num_samples))),
        dict(x=np.ones((1, g.num_samples), dtype=np.int32), init="test"),
        dict(x=np.ones((7, g.num_samples)), init=np.ones(8)),
    ]
    for kw in bad:
        with pytest.raises((ValueError, TypeError)):
            grg.matmul(kw["x"], "up", init=kw["init"])


@pytest.mark.parametrize("mode", ["all_streamed", "partial"])
def test_streamed_layouts_on_fixture(fixtures, mode):
    """The reference's GRGs with blocks streamed through the pinned ring (every block,
    or the part a budget between the streamed and the fully resident sizes leaves)."""
    from pygrgl_spmv.backends.hip import HipRuntime

    path, g, X, Miss = fixtures["t200"]
    req = full_requirements(max_k_up=4, max_k_down=4)
    if mode == "all_streamed":
        layout = build_hip_layout([path], requirements=req, ring_buffer_size=2, allow_residency=False)
    else:
        full = build_hip_layout([path], requirements=req, vram_budget_bytes=0)
        streamed = build_hip_layout([path], requirements=req, ring_buffer_size=2, allow_residency=False)
        budget = (streamed.bytes_total + full.required_budget_for_full_residency) // 2
        layout = build_hip_layout([path], requirements=req, vram_budget_bytes=budget, ring_buffer_size=2)
    flags = [b.resident for a in layout.artifacts for b in (*a.blocks_up, *a.blocks_down)]
    assert not all(flags) and (mode == "all_streamed") == (not any(flags))
    rng = np.random.default_rng(55)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        x = rng.random((4, g.num_samples))
        np.testing.assert_allclose(grg.matmul(x, "up"), x @ X, rtol=1e-12, atol=1e-12)
        y, m = rng.random((4, g.num_mutations)), rng.random((4, g.num_mutations))
        np.testing.assert_allclose(grg.matmul(y, "down", miss=m), y @ X.T + m @ Miss.T, rtol=1e-12, atol=1e-12)
        D = _diploid(X)
        ones = np.ones((2, g.num_samples))
        np.testing.assert_array_equal(grg.matmul(ones, "up", init="xtx"), np.tile((D * D).sum(axis=0), (2, 1)))


@pytest.mark.parametrize("name", ["t200", "msprime"])
def test_fp32c_on_fixture(fixtures, name):
    """The FP32-compensated form on the reference's GRGs: exact for +-1 inputs (integer
    sums below 2^24), within 5e-7 normwise for random inputs, missing data included."""
    from pygrgl_spmv.backends.hip import HipRuntime

    path, g, X, Miss = fixtures[name]
    rng = np.random.default_rng(77)
    layout = build_hip_layout([path], dtype=np.float32, requirements=full_requirements(max_k_up=8, max_k_down=8))
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        s = rng.choice(np.array([-1.0, 1.0]), size=(8, g.num_samples))
        np.testing.assert_array_equal(grg.matmul(s.astype(np.float32), "up"), (s @ X).astype(np.float32))
        t = rng.choice(np.array([-1.0, 1.0]), size=(8, g.num_mutations))
        np.testing.assert_array_equal(grg.matmul(t.astype(np.float32), "down"), (t @ X.T).astype(np.float32))
        y, m = rng.standard_normal((4, g.num_mutations)), rng.standard_normal((4, g.num_mutations))
        got = grg.matmul(y.astype(np.float32), "down", miss=m.astype(np.float32))
        assert O.normwise_rel_err(got, y @ X.T + m @ Miss.T) <= 5e-7
        x = rng.standard_normal((4, g.num_individuals))
        got = grg.matmul(x.astype(np.float32), "up", by_individual=True)
        assert O.normwise_rel_err(got, x @ _diploid(X)) <= 5e-7
