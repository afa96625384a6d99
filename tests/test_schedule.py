"""Level-launch schedules of the HIP runtime (CPU): heavy rows and split rows.

The kernel's partials launch gives item i of heavy row h the entries
[j * chunk, (j + 1) * chunk) of the row's concatenated segments, with
j = i - split_ptr[h]; these tests check that the items of every split row cover
its degree exactly once and that unsplit rows have no items.
"""

from __future__ import annotations

import numpy as np
import pytest

from pygrgl_spmv.backends.hip.runtime import column_shapes, launch_shape, level_schedule


def _table(sched, table, cursor):
    heavy = table[sched.heavy_offset - cursor: sched.heavy_offset - cursor + sched.num_heavy]
    if not sched.num_items:
        return heavy, None, None
    ptr = table[sched.split_ptr_offset - cursor: sched.split_ptr_offset - cursor + sched.num_heavy + 1]
    rows = table[sched.split_row_offset - cursor: sched.split_row_offset - cursor + sched.num_items]
    return heavy, ptr, rows


def test_heavy_rows_only_above_threshold():
    deg = np.tile(np.array([1, 2, 3, 0], dtype=np.int64), 100)
    deg[[2, 5, 6]] = [600, 513, 512]  # mean stays below 64, so the floor 512 rules
    sched, table = level_schedule(deg, 0, deg.size, heavy_min=512, cursor=7)
    heavy, ptr, _ = _table(sched, table, 7)
    assert sched.heavy_threshold == 512
    np.testing.assert_array_equal(heavy, [2, 5])
    assert ptr is None and sched.num_items == 0
    light = deg[deg <= 512]
    assert sched.mean_degree == pytest.approx(max(1.0, light.mean()))


def test_heavy_threshold_follows_level_mean():
    deg = np.full(100, 200, dtype=np.int64)
    deg[7] = 1700  # mean 215 -> threshold 8 x mean = 1720 keeps row 7 light
    sched, _ = level_schedule(deg, 0, deg.size, heavy_min=512, cursor=0)
    assert sched.heavy_threshold == int(8 * deg.mean())
    assert sched.num_heavy == 0


@pytest.mark.parametrize("chunk", [1, 5, 64])
def test_split_items_cover_each_row_once(chunk):
    rng = np.random.default_rng(4)
    deg = rng.integers(0, 40, size=500)
    deg[[3, 77, 78, 400]] = [5000, 700, 12_345, 2049]
    lo, hi = 2, 480
    sched, table = level_schedule(deg, lo, hi, heavy_min=512, cursor=11, split_min=2048, split_chunk=chunk)
    heavy, ptr, rows = _table(sched, table, 11)
    np.testing.assert_array_equal(heavy, [3, 77, 78, 400])
    split = deg[heavy] > 2048
    assert ptr[0] == 0 and np.all(np.diff(ptr) >= 0)
    for h, row in enumerate(heavy):
        n_items = int(ptr[h + 1] - ptr[h])
        if not split[h]:
            assert n_items == 0
            continue
        assert n_items == -(-int(deg[row]) // chunk)
        covered = sum(min(chunk, int(deg[row]) - j * chunk) for j in range(n_items))
        assert covered == deg[row]
        np.testing.assert_array_equal(rows[ptr[h]: ptr[h + 1]], h)
    assert sched.num_items == int(ptr[-1]) == rows.size
    assert table.dtype == np.int32 and table.size == 2 * sched.num_heavy + 1 + sched.num_items


def test_no_split_when_disabled():
    deg = np.ones(1000, dtype=np.int64)
    deg[500] = 100_000
    sched, table = level_schedule(deg, 0, deg.size, heavy_min=512, cursor=0)
    assert sched.num_heavy == 1 and sched.num_items == 0 and table.size == 1


def test_rows_to_split_are_heavy_even_in_a_level_of_long_rows():
    """A level made only of very long rows (mean above split_min) still splits them."""
    deg = np.full(8, 30_000, dtype=np.int64)
    sched, table = level_schedule(deg, 0, deg.size, heavy_min=512, cursor=0, split_min=4096, split_chunk=2048)
    heavy, ptr, _ = _table(sched, table, 0)
    assert sched.num_heavy == 8 and sched.heavy_threshold == 4096
    np.testing.assert_array_equal(np.diff(ptr), 15)


def test_column_shapes_respect_alignment():
    from pygrgl_spmv.backends.hip import ffi

    assert column_shapes(ffi.F64, 16, [16, 16]) == [(4, 4), (2, 8), (1, 16)]
    assert column_shapes(ffi.F64, 6, [6]) == [(2, 4), (1, 8)]
    assert column_shapes(ffi.F32C, 3, [3]) == [(1, 4)]
    # a leading dimension that is not a multiple of the vector width excludes it
    assert (4, 4) not in column_shapes(ffi.F64, 16, [16, 18])
    cpl, kl, es = launch_shape(ffi.F64, 16, [16], mean_degree=40.0)
    assert (cpl, kl) == (2, 8) and es == 4 and kl * es <= 64
