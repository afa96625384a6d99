"""Row groups of the MFMA row-group launch (host builder, no GPU)."""

from __future__ import annotations

import numpy as np
import pytest

from pygrgl_spmv.backends.hip.host import GROUP_ROWS, build_groups


def _random_rows(rng, n_rows, max_deg, n_src, segments=1):
    deg = rng.integers(0, max_deg + 1, size=n_rows)
    eptr = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(deg, out=eptr[1:])
    keys = []
    for d in deg:
        src = rng.choice(n_src, size=int(d), replace=False).astype(np.uint32)
        seg = rng.integers(0, segments, size=int(d)).astype(np.uint32)
        keys.append(np.sort((seg << 30) | src))
    return eptr.astype(np.int32), (np.concatenate(keys) if keys else np.zeros(0, np.uint32)).astype(np.uint32), deg


@pytest.mark.parametrize("heavy", [-1, 9])
@pytest.mark.parametrize("lo,hi", [(0, 200), (5, 187), (13, 13)])
def test_groups_match_brute_force(lo, hi, heavy):
    rng = np.random.default_rng(40 + lo + hi + heavy)
    eptr, keys, deg = _random_rows(rng, 200, 14, 60, segments=3)
    grp = build_groups(eptr, keys, lo, hi, heavy_threshold=heavy, threads=3)
    n_groups = -(-(hi - lo) // GROUP_ROWS)
    assert grp.num_groups == n_groups and grp.off.size == n_groups + 1 and grp.off[0] == 0
    for g in range(n_groups):
        r0, r1 = lo + GROUP_ROWS * g, min(hi, lo + GROUP_ROWS * (g + 1))
        light = [r for r in range(r0, r1) if heavy < 0 or deg[r] <= heavy]
        want_store = sum(1 << (r - r0) for r in light)
        assert int(grp.store[g]) == want_store
        masks = {}
        for r in light:
            for k in keys[eptr[r]:eptr[r + 1]]:
                masks[int(k)] = masks.get(int(k), 0) | (1 << (r - r0))
        e0, e1 = int(grp.off[g]), int(grp.off[g + 1])
        assert (e1 - e0) % 4 == 0 and e1 - e0 - len(masks) in (0, 1, 2, 3)
        got_keys = grp.key[e0:e1]
        got_masks = grp.mask[e0:e1]
        n = len(masks)
        assert got_keys[:n].tolist() == sorted(masks)
        assert got_masks[:n].tolist() == [masks[k] for k in sorted(masks)]
        assert np.all(got_keys[n:] == 0xFFFFFFFF) and np.all(got_masks[n:] == 0)


def test_group_entries_reproduce_row_sums():
    """Summing each group entry's source value into the rows of its mask gives every
    light row's sum over its own entries (the row-group launch's arithmetic)."""
    rng = np.random.default_rng(7)
    eptr, keys, deg = _random_rows(rng, 300, 20, 120)
    values = rng.integers(-50, 50, size=(120, 5)).astype(np.float64)
    grp = build_groups(eptr, keys, 0, 300)
    got = np.zeros((300, 5))
    for g in range(grp.num_groups):
        for e in range(int(grp.off[g]), int(grp.off[g + 1])):
            k, m = int(grp.key[e]), int(grp.mask[e])
            if k == 0xFFFFFFFF:
                continue
            for i in range(GROUP_ROWS):
                if m >> i & 1:
                    got[GROUP_ROWS * g + i] += values[k]
    want = np.stack([values[keys[eptr[r]:eptr[r + 1]]].sum(axis=0) for r in range(300)])
    np.testing.assert_array_equal(got, want)
    assert grp.num_entries < int(eptr[-1])  # neighbouring rows share sources


def test_repeated_key_in_a_row_is_refused():
    eptr = np.asarray([0, 3], dtype=np.int32)
    keys = np.asarray([4, 4, 9], dtype=np.uint32)
    with pytest.raises(RuntimeError, match="same entry key twice"):
        build_groups(eptr, keys, 0, 1)
