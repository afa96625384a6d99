"""Host tile builder (``grg_host_build_tiles``): structure invariants, and the
tiled row sums restated on the CPU equal the plain CSR row sums."""

from __future__ import annotations

import numpy as np
import pytest

from pygrgl_spmv.backends.hip.host import build_tiles


def _random_csr(rng, rows, src_rows, max_deg, heavy_rows=()):
    deg = rng.integers(0, max_deg + 1, size=rows)
    for r, d in heavy_rows:
        deg[r] = d
    deg = np.minimum(deg, src_rows)
    ptr = np.zeros(rows + 1, dtype=np.int64)
    np.cumsum(deg, out=ptr[1:])
    idx = np.concatenate([np.sort(rng.choice(src_rows, size=int(d), replace=False)) for d in deg] or [np.zeros(0, int)])
    return ptr.astype(np.int32), idx.astype(np.int32)


def _tiled_sums(t, segs, values, off, heavy):
    """Row sums through the tiles, as the tiled launch computes them (heavy rows left at 0)."""
    R = int(off[-1])
    k = values[0].shape[1]
    out = np.zeros((R, k))
    for ti in range(int(t.level_tile_off[-1])):
        r0, r1 = int(t.tile_row[ti]), int(t.tile_row[ti + 1])
        s0, s1 = int(t.tile_src[ti]), int(t.tile_src[ti + 1])
        keys = t.src_list[s0:s1]
        staged = np.stack([values[int(kk >> 30)][int(kk & 0x3FFFFFFF)] for kk in keys]) if s1 > s0 else None
        for r in range(r0, r1):
            lvl = int(np.searchsorted(off, r, side="right") - 1)
            deg = int(t.eptr[r + 1] - t.eptr[r])
            if heavy[lvl] >= 0 and deg > heavy[lvl]:
                continue
            if s1 == s0:  # direct tile
                for (p, i, _), v in zip(segs, values):
                    out[r] += v[i[p[r]:p[r + 1]]].sum(axis=0)
            else:
                out[r] = staged[t.lidx[t.eptr[r]:t.eptr[r + 1]]].sum(axis=0)
    return out


@pytest.mark.parametrize("max_staged,max_rows,max_entries", [(8, 64, 4096), (32, 16, 4096), (512, 1024, 4096), (64, 64, 30)])
def test_tiles_reproduce_row_sums(max_staged, max_rows, max_entries):
    rng = np.random.default_rng(7)
    off = np.array([0, 40, 300, 420, 421, 500])
    R = int(off[-1])
    L = off.size - 1
    segs = [(*_random_csr(rng, R, 300, 12, heavy_rows=[(50, 40), (420, 25)]), 300),
            (*_random_csr(rng, R, 90, 3), 90)]
    heavy = np.array([-1, 30, -1, 20, -1])
    t = build_tiles(off, segs, heavy, max_staged=max_staged, max_rows=max_rows, max_entries=max_entries)
    # tiles cover every row in order, levels start on tile boundaries
    assert t.tile_row[0] == 0 and t.tile_row[-1] == R and np.all(np.diff(t.tile_row) > 0)
    for l in range(L):
        assert t.tile_row[t.level_tile_off[l]] == off[l] or off[l] == off[l + 1]
    assert np.all(np.diff(t.tile_src) >= 0)
    staged = np.diff(t.tile_src)
    assert staged.max() <= max_staged
    assert np.all(np.diff(t.tile_row) <= max_rows)
    rows = np.diff(t.tile_row)
    ents = t.eptr[t.tile_row[1:]] - t.eptr[t.tile_row[:-1]]
    assert np.all(ents[staged > 0] <= max_entries)
    for l in range(L):
        lo, hi = t.level_tile_off[l], t.level_tile_off[l + 1]
        st = staged[lo:hi] > 0
        assert t.level_max_staged[l] == (staged[lo:hi].max() if hi > lo else 0)
        assert t.level_max_rows[l] == (rows[lo:hi][st].max() if st.any() else 0)
        assert t.level_max_entries[l] == (ents[lo:hi][st].max() if st.any() else 0)
    # staged sources are distinct within a tile
    for ti in range(int(t.level_tile_off[-1])):
        keys = t.src_list[t.tile_src[ti]:t.tile_src[ti + 1]]
        assert np.unique(keys).size == keys.size
    values = [rng.standard_normal((300, 3)), rng.standard_normal((90, 3))]
    want = np.zeros((R, 3))
    for (p, i, _), v in zip(segs, values):
        want += np.stack([v[i[p[r]:p[r + 1]]].sum(axis=0) for r in range(R)])
    deg = np.diff(t.eptr)
    lvl = np.searchsorted(off, np.arange(R), side="right") - 1
    is_heavy = (heavy[lvl] >= 0) & (deg > heavy[lvl])
    assert is_heavy[50] and is_heavy[420]
    # heavy rows sit alone in direct tiles
    for r in np.flatnonzero(is_heavy):
        ti = int(np.searchsorted(t.tile_row, r, side="right") - 1)
        assert t.tile_row[ti] == r and t.tile_row[ti + 1] == r + 1 and staged[ti] == 0
    want[is_heavy] = 0.0
    np.testing.assert_allclose(_tiled_sums(t, segs, values, off, heavy), want, rtol=1e-13, atol=1e-13)


def test_long_rows_get_direct_tiles():
    rng = np.random.default_rng(3)
    off = np.array([0, 64])
    p, i = _random_csr(rng, 64, 200, 4, heavy_rows=[(10, 100), (11, 120)])
    t = build_tiles(off, [(p, i, 200)], np.array([-1]), max_staged=50, max_rows=1024)
    starts = list(t.tile_row[:-1])
    for r in (10, 11):
        ti = starts.index(r)
        assert t.tile_row[ti + 1] == r + 1 and t.tile_src[ti + 1] == t.tile_src[ti]


def test_builder_rejects_bad_input():
    off = np.array([0, 4])
    p = np.array([0, 1, 2, 3, 4], dtype=np.int32)
    i = np.array([0, 1, 2, 9], dtype=np.int32)
    with pytest.raises(RuntimeError, match="out of range"):
        build_tiles(off, [(p, i, 5)], np.array([-1]))
    with pytest.raises(ValueError, match="cover every row"):
        build_tiles(off, [(p[:3], i, 10)], np.array([-1]))


def test_merged_keys_follow_segment_order():
    from pygrgl_spmv.backends.hip.host import merged_keys

    rng = np.random.default_rng(11)
    R = 50
    segs = [(*_random_csr(rng, R, 40, 5), 40), (*_random_csr(rng, R, 7, 2), 7), (*_random_csr(rng, R, 9, 3), 9)]
    deg = sum(np.diff(p.astype(np.int64)) for p, _, _ in segs)
    eptr = np.zeros(R + 1, dtype=np.int64)
    np.cumsum(deg, out=eptr[1:])
    keys = merged_keys(eptr.astype(np.int32), segs)
    for r in range(R):
        want = np.concatenate([(np.uint32(s) << np.uint32(30)) | i[p[r]:p[r + 1]].astype(np.uint32)
                               for s, (p, i, _) in enumerate(segs)])
        np.testing.assert_array_equal(keys[eptr[r]:eptr[r + 1]], want)
    with pytest.raises(RuntimeError, match="eptr"):
        merged_keys((eptr + 1).astype(np.int32), segs)


PAD = 255  # entry position of the padding entries


def _panel_sums(pn, values, row_begin, row_end):
    """Row sums through the panels, pass by pass, as the panel launch computes them."""
    R = pn.panel_rows
    k = values[0].shape[1]
    out = np.zeros((row_end - row_begin, k))
    for p in range(pn.num_panels):
        for g in range(int(pn.pan_pass[p]), int(pn.pan_pass[p + 1])):
            keys = pn.src_list[pn.pass_src[g]:pn.pass_src[g + 1]]
            assert 0 < keys.size <= pn.max_staged
            staged = np.stack([values[int(kk >> 30)][int(kk & 0x3FFFFFFF)] for kk in keys])
            off = pn.pass_rowoff[g * (R + 1):(g + 1) * (R + 1)].astype(np.int64)
            ent = pn.lidx[pn.pass_ent[g]:pn.pass_ent[g + 1]].astype(np.int64)
            assert off[0] == 0 and off[-1] == ent.size and np.all(np.diff(off) >= 0)
            assert np.all(off % 4 == 0) and int(pn.pass_ent[g]) % 4 == 0  # runs padded to multiples of 4
            for i in range(R):
                r = p * R + i
                run = ent[off[i]:off[i + 1]]
                pos = run[run != PAD]
                assert np.all(run[pos.size:] == PAD) and run.size - pos.size < 4  # padding at the end, < 4
                assert np.all(np.diff(pos) > 0) and np.all(pos < keys.size)  # ascending, each source once
                if r < row_end - row_begin:
                    out[r] += staged[pos].sum(axis=0)
                else:
                    assert pos.size == 0
    return out


@pytest.mark.parametrize("panel_rows,max_staged,max_entries,threads",
                         [(64, 128, 8000, 1), (16, 7, 65535, 3), (1, 255, 4, 2), (5, 1, 20, 4), (1024, 255, 4096, 2),
                          (256, 255, 1024, 3), (7, 255, 40, 1)])
def test_panels_reproduce_row_sums(panel_rows, max_staged, max_entries, threads):
    from pygrgl_spmv.backends.hip.host import build_panels, merged_keys

    rng = np.random.default_rng(11)
    rows = 300
    segs = [(*_random_csr(rng, rows, 500, 40, heavy_rows=[(7, 200)]), 500), (*_random_csr(rng, rows, 60, 4), 60)]
    deg = sum(np.diff(p.astype(np.int64)) for p, _, _ in segs)
    eptr = np.zeros(rows + 1, dtype=np.int32)
    np.cumsum(deg, out=eptr[1:])
    keys = merged_keys(eptr, segs)
    values = [rng.integers(-8, 9, size=(n, 3)).astype(np.float64) for _, _, n in segs]
    plain = np.zeros((rows, 3))
    for (p, i, _), v in zip(segs, values):
        for r in range(rows):
            plain[r] += v[i[p[r]:p[r + 1]]].sum(axis=0)
    heavy = 100
    for lo, hi in ((0, rows), (13, 250), (40, 40)):
        pn = build_panels(eptr, keys, lo, hi, panel_rows=panel_rows, max_staged=max_staged,
                          max_pass_entries=max_entries, heavy_threshold=heavy, threads=threads)
        assert pn.num_panels == -(-(hi - lo) // panel_rows)
        assert pn.max_passes == (int(np.diff(pn.pan_pass).max()) if pn.num_panels else 0)
        # passes are greedy runs of the panel's sources: each holds at most max_staged
        # sources and max_entries entries, and the next source would break a limit
        for p in range(pn.num_panels):
            g0, g1 = int(pn.pan_pass[p]), int(pn.pan_pass[p + 1])
            n_src = np.diff(pn.pass_src[g0:g1 + 1])
            n_ent = np.diff(pn.pass_ent[g0:g1 + 1])
            assert np.all((n_src > 0) & (n_src <= max_staged)) and np.all(n_ent <= max_entries)
            keys_p = pn.src_list[pn.pass_src[g0]:pn.pass_src[g1]]
            assert np.all(np.diff(keys_p.astype(np.int64)) > 0)  # distinct, ascending over the panel
            # the next pass's first source would have broken a limit: its rows' padded runs grow
            for j, g in enumerate(range(g0, g1 - 1)):
                off = pn.pass_rowoff[g * (panel_rows + 1):(g + 1) * (panel_rows + 1)].astype(np.int64)
                ent = pn.lidx[pn.pass_ent[g]:pn.pass_ent[g + 1]]
                real = np.array([int(np.count_nonzero(ent[off[i]:off[i + 1]] != PAD)) for i in range(panel_rows)])
                nxt = pn.src_list[pn.pass_src[g + 1]]
                g2 = g + 1
                off2 = pn.pass_rowoff[g2 * (panel_rows + 1):(g2 + 1) * (panel_rows + 1)].astype(np.int64)
                ent2 = pn.lidx[pn.pass_ent[g2]:pn.pass_ent[g2 + 1]]
                rows_nxt = [i for i in range(panel_rows) if 0 in ent2[off2[i]:off2[i + 1]]]
                grow = sum(4 for i in rows_nxt if real[i] % 4 == 0)
                assert n_src[j] == max_staged or n_ent[j] + grow > max_entries
                del nxt
        got = _panel_sums(pn, values, lo, hi)
        want = plain[lo:hi].copy()
        want[deg[lo:hi] > heavy] = 0  # heavy rows are left to their heavy workgroups
        np.testing.assert_array_equal(got, want)
        # every non-heavy entry appears exactly once
        assert int(np.count_nonzero(pn.lidx[:int(pn.pass_ent[-1])] != PAD)) == int(deg[lo:hi][deg[lo:hi] <= heavy].sum())


def test_panels_reject_bad_arguments():
    from pygrgl_spmv.backends.hip.host import build_panels

    eptr = np.array([0, 2, 2], dtype=np.int32)
    keys = np.array([3, 3], dtype=np.uint32)  # a row listing one source twice
    with pytest.raises(RuntimeError, match="twice"):
        build_panels(eptr, keys, 0, 2, panel_rows=2, max_staged=4)
    with pytest.raises(RuntimeError, match="panel_rows"):
        build_panels(eptr, np.array([1, 3], dtype=np.uint32), 0, 2, panel_rows=0, max_staged=4)
    with pytest.raises(RuntimeError, match="panel_rows"):
        build_panels(eptr, np.array([1, 3], dtype=np.uint32), 0, 2, panel_rows=2048, max_staged=4)
    with pytest.raises(RuntimeError, match="max_pass_entries"):
        build_panels(eptr, np.array([1, 3], dtype=np.uint32), 0, 2, panel_rows=8, max_staged=4, max_pass_entries=31)
    with pytest.raises(RuntimeError, match="max_staged"):
        build_panels(eptr, np.array([1, 3], dtype=np.uint32), 0, 2, panel_rows=8, max_staged=256)
