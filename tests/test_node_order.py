"""Device node order (``grg_host_locality_order``): a permutation inside every
level that keeps samples as the prefix of level 0, and, with degree windows,
orders each run of consecutive nodes by child count without moving any node
out of its run."""

from __future__ import annotations

import numpy as np
import pytest

from pygrgl_spmv.backends.hip.host import locality_order
from pygrgl_spmv.grg.artifact import iter_artifact_blocks, load_grg_spmv
from pygrgl_spmv.synth import SynthConfig, write_artifact


@pytest.fixture(scope="module")
def dag(artifact_dir):
    cfg = SynthConfig(num_samples=600, num_mutations=6000, seed=4242, num_trees=300, keep_fraction=0.3,
                      window_fraction=0.1, num_segments=2)
    path = write_artifact(cfg, artifact_dir / "order.grg_spmv")
    st = load_grg_spmv(path, np.float64)
    off = np.asarray(st.level_offsets, dtype=np.int64)
    blocks = [(b.dst_level, b.src_level, np.asarray(b.indptr, np.int64), np.asarray(b.indices, np.int32))
              for b in iter_artifact_blocks(path) if b.nnz]
    deg = np.zeros(st.num_nodes, dtype=np.int64)
    for d, _, ptr, _ in blocks:
        deg[off[d]:off[d + 1]] += np.diff(ptr)
    return st.num_nodes, st.num_samples, off, blocks, deg


def _check_permutation(new_pos, N, S, off):
    assert np.array_equal(np.sort(new_pos), np.arange(N))
    lvl_old = np.searchsorted(off, np.arange(N), side="right") - 1
    lvl_new = np.searchsorted(off, new_pos, side="right") - 1
    assert np.array_equal(lvl_old, lvl_new)
    assert np.array_equal(np.sort(new_pos[:S]), np.arange(S))


@pytest.mark.parametrize("window,min_mean", [(4, 0.0), (64, 0.0), (16, 3.0)])
def test_degree_windows(dag, window, min_mean):
    N, S, off, blocks, deg = dag
    plain = locality_order(N, S, off, blocks)
    win = locality_order(N, S, off, blocks, degree_window=window, degree_min_mean=min_mean)
    _check_permutation(plain, N, S, off)
    _check_permutation(win, N, S, off)
    deg_plain = np.empty(N, dtype=np.int64)
    deg_plain[plain] = deg
    deg_win = np.empty(N, dtype=np.int64)
    deg_win[win] = deg
    node_plain = np.empty(N, dtype=np.int64)
    node_plain[plain] = np.arange(N)
    node_win = np.empty(N, dtype=np.int64)
    node_win[win] = np.arange(N)
    sorted_levels = 0
    for l in range(off.size - 1):
        lo, hi = int(off[l]), int(off[l + 1])
        if l == 0 or hi == lo or deg[lo:hi].mean() < min_mean:
            # untouched level (level 0 always): identical to the plain order
            assert np.array_equal(node_win[lo:hi], node_plain[lo:hi])
            continue
        sorted_levels += 1
        for w in range(lo, hi, window):
            we = min(hi, w + window)
            # same nodes in the run, ordered by child count, ties in plain order
            assert set(node_win[w:we]) == set(node_plain[w:we])
            assert np.all(np.diff(deg_win[w:we]) >= 0)
            rank = {int(v): i for i, v in enumerate(node_plain[w:we])}
            d = deg_win[w:we]
            r = np.array([rank[int(v)] for v in node_win[w:we]])
            same = d[1:] == d[:-1]
            assert np.all(r[1:][same] > r[:-1][same])
    assert sorted_levels > 0


def test_degree_window_rejects_negative(dag):
    N, S, off, blocks, _ = dag
    with pytest.raises(RuntimeError, match="degree_window"):
        locality_order(N, S, off, blocks, degree_window=-1)


def _first_child_pos(N, off, blocks, pos):
    """Position (under ``pos``) of each node's first child, -1 for childless nodes."""
    mn = np.full(N, np.iinfo(np.int64).max)
    for d, s, ptr, idx in blocks:
        rows = np.repeat(np.arange(ptr.size - 1, dtype=np.int64), np.diff(ptr)) + off[d]
        np.minimum.at(mn, rows, pos[idx.astype(np.int64) + off[s]])
    return mn


@pytest.mark.parametrize("window,degree_window,min_rows", [(8, 0, 0), (64, 0, 0), (64, 8, 0), (32, 4, 40)])
def test_cluster_order(dag, window, degree_window, min_rows):
    """Min-child order: each window of a refined level holds the same nodes as before, ordered
    by the position of their first child (ties in the previous order); aligned degree runs inside
    it keep their node sets and are ordered by child count; other levels are untouched."""
    from pygrgl_spmv.backends.hip.host import cluster_order

    N, S, off, blocks, deg = dag
    base = locality_order(N, S, off, blocks)
    got = cluster_order(base, off, blocks, window=window, degree_window=degree_window, min_mean=0.0, min_rows=min_rows)
    _check_permutation(got, N, S, off)
    assert np.array_equal(cluster_order(base, off, blocks, window=window, degree_window=degree_window, min_mean=0.0,
                                        min_rows=min_rows), got)  # deterministic
    node_base = np.empty(N, dtype=np.int64)
    node_base[base] = np.arange(N)
    node_got = np.empty(N, dtype=np.int64)
    node_got[got] = np.arange(N)
    first = _first_child_pos(N, off, blocks, got)  # children positions are final (lower levels)
    refined = 0
    for l in range(off.size - 1):
        lo, hi = int(off[l]), int(off[l + 1])
        if l == 0 or hi - lo < max(2, min_rows):
            assert np.array_equal(node_got[lo:hi], node_base[lo:hi])
            continue
        refined += 1
        for w in range(lo, hi, window):
            we = min(hi, w + window)
            assert set(node_got[w:we]) == set(node_base[w:we])
            # the expected window, restated: stable sort by first-child position, then degree runs
            exp = sorted(node_base[w:we].tolist(), key=lambda v: first[v])
            if degree_window > 1:
                for u in range(0, len(exp), degree_window):
                    exp[u:u + degree_window] = sorted(exp[u:u + degree_window], key=lambda v: deg[v])
            assert node_got[w:we].tolist() == exp
            if degree_window <= 1:
                assert np.all(np.diff(first[node_got[w:we]]) >= 0)
    assert refined > 0


def test_cluster_order_rejects_bad_arguments(dag):
    from pygrgl_spmv.backends.hip.host import cluster_order

    N, S, off, blocks, _ = dag
    base = locality_order(N, S, off, blocks)
    with pytest.raises(RuntimeError, match="window must be >= 2"):
        cluster_order(base, off, blocks, window=1)
    with pytest.raises(RuntimeError, match="degree_window"):
        cluster_order(base, off, blocks, window=8, degree_window=-1)
    bad = base.copy()
    bad[0] = N
    with pytest.raises(RuntimeError, match="new_pos out of range"):
        cluster_order(bad, off, blocks, window=8)
    # a node moved to another level's positions (a level 2 node given a level 1 position)
    lo1, lo2 = int(off[1]), int(off[2])
    swapped = base.copy()
    a, b = int(np.flatnonzero(base == lo1)[0]), int(np.flatnonzero(base == lo2)[0])
    swapped[a], swapped[b] = base[b], base[a]
    with pytest.raises(RuntimeError, match="not a permutation inside each level"):
        cluster_order(swapped, off, blocks, window=8)
