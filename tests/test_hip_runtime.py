"""HIP runtime behaviour on the GPU: split rows, untuned launches, streamed ring
layouts, hipGraph capture, several GRGs in one runtime, lifecycle.

Counterparts in the reference: streaming and budget cases of
``runtime/test_triton.py:230-576`` and ``_streaming_cases.py``, graph capture
``test_triton.py:920-1058``, runtime protocol ``runtime/test_prepare_cuda.py``.
Every numeric check compares against the CPU oracle on the same inputs: FP64
within 1e-12 normwise relative error, exactly for +-1 / integer inputs.
"""

from __future__ import annotations

from contextlib import ExitStack

import numpy as np
import pytest

from oracle import grg_oracle as O
from tests._builders import build_hip_layout, full_requirements, write_band_artifact, write_synthetic_artifact

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 5e-7   # FP32-compensated sweeps: the FP32 storage floor (measured 1e-7 to 3e-7)


def _cfg(seed, samples=300, mutations=2000, ploidy=2, missing=0.05):
    from pygrgl_spmv.synth import SynthConfig

    return SynthConfig(num_samples=samples, num_mutations=mutations, seed=seed, ploidy=ploidy, num_trees=120,
                       keep_fraction=0.3, window_fraction=0.1, num_segments=3, missing_fraction=missing, missing_pool=6)


@pytest.fixture(scope="module")
def art_a(artifact_dir):
    return write_synthetic_artifact(artifact_dir, "runtime-a", _cfg(201))


@pytest.fixture(scope="module")
def art_b(artifact_dir):
    return write_synthetic_artifact(artifact_dir, "runtime-b", _cfg(202, samples=180, mutations=900, ploidy=1, missing=0.0))


@pytest.fixture(scope="module")
def oracle_a(art_a):
    return O.load_artifact(art_a)


def _cols(g, direction, by_individual=False):
    if direction == "up":
        return g.num_individuals if by_individual else g.num_samples
    return g.num_mutations


def _check(got, exp):
    assert got.shape == exp.shape
    assert O.normwise_rel_err(got, exp) <= FP64_TOL


# ---------------------------------------------------------------------------
# split rows (partials launch + level launch)
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def split_runtime(art_a):
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=8, max_k_down=8))
    with HipRuntime(layout, split_min=12, split_chunk=5) as rt:
        yield rt


def test_split_schedule_present(split_runtime):
    art = split_runtime._artifacts[0]
    items = [s.num_items for s in art.sched.values()]
    assert max(items) > 0, "the small split parameters must split some rows"


@pytest.mark.parametrize("direction", ["up", "down"])
@pytest.mark.parametrize("init_mode", ["none", "xtx", "matrix"])
def test_split_rows_match_oracle(split_runtime, oracle_a, direction, init_mode):
    (grg,) = split_runtime.grgs
    rng = np.random.default_rng(3100 + len(init_mode) + (7 if direction == "down" else 0))
    x = rng.standard_normal((3, _cols(oracle_a, direction)))
    init = {"none": None, "xtx": "xtx", "matrix": rng.standard_normal((3, oracle_a.num_nodes))}[init_mode]
    _check(grg.matmul(x, direction, init=init), O.matmul(oracle_a, x, direction, init=init))


def test_split_rows_missing_down_and_exact(split_runtime, oracle_a):
    (grg,) = split_runtime.grgs
    rng = np.random.default_rng(3200)
    y = rng.choice(np.array([-1.0, 1.0]), size=(5, oracle_a.num_mutations))
    miss = rng.choice(np.array([-1.0, 1.0]), size=(5, oracle_a.num_mutations))
    np.testing.assert_array_equal(grg.matmul(y, "down", miss=miss.copy()), O.matmul(oracle_a, y, "down", miss=miss.copy()))
    x = rng.choice(np.array([-1.0, 1.0]), size=(8, oracle_a.num_individuals))
    np.testing.assert_array_equal(grg.matmul(x, "up", by_individual=True), O.matmul(oracle_a, x, "up", by_individual=True))


# ---------------------------------------------------------------------------
# default launch shapes (no autotuning)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("k", [1, 4, 9])
def test_untuned_launch_shapes(art_a, oracle_a, k):
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=16, max_k_down=16))
    with HipRuntime(layout, autotune=False) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3300 + k)
        for direction in ("up", "down"):
            x = rng.standard_normal((k, _cols(oracle_a, direction)))
            _check(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("unroll", [4, 8])
@pytest.mark.parametrize("es", [1, 4, 16])
def test_streamed_light_path(art_a, oracle_a, dtype, unroll, es):
    """The software-pipelined light path (waves looping over row batches, small and
    capped grids) on every level and write-out launch it applies to, with merged
    entry keys for the multi-segment down levels and heavy / split rows beside it:
    exact for +-1 inputs, FP64 within 1e-12, FP32c within 5e-7."""
    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.backends.hip import ffi

    layout = build_hip_layout([art_a], dtype=dtype, requirements=full_requirements(max_k_up=16, max_k_down=16))
    shape = {"unroll": unroll, "rows_per_group": 1, "es": es, "flags": ffi.FLAG_STREAM}
    tol = FP64_TOL if dtype == np.float64 else FP32_TOL
    with HipRuntime(layout, fixed_shape=shape, split_min=40, split_chunk=16) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3377 + unroll + es)
        for k in (2, 16):
            for direction in ("up", "down"):
                x = rng.choice(np.array([-1.0, 1.0]), size=(k, _cols(oracle_a, direction))).astype(dtype)
                np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x.astype(np.float64), direction))
        y = rng.standard_normal((5, oracle_a.num_mutations))
        miss = rng.standard_normal((5, oracle_a.num_mutations))
        got = grg.matmul(y.astype(dtype), "down", miss=miss.astype(dtype))
        assert O.normwise_rel_err(got, O.matmul(oracle_a, y, "down", miss=miss)) <= tol
        x = rng.standard_normal((8, oracle_a.num_individuals))
        got = grg.matmul(x.astype(dtype), "up", by_individual=True, init="xtx")
        assert O.normwise_rel_err(got, O.matmul(oracle_a, x, "up", by_individual=True, init="xtx")) <= tol
    # a one-workgroup grid: every wave loops over many batches
    shape["stream_blocks"] = 1
    with HipRuntime(layout, fixed_shape=shape) as rt:
        (grg,) = rt.grgs
        for direction in ("up", "down"):
            x = rng.standard_normal((3, _cols(oracle_a, direction))).astype(dtype)
            assert O.normwise_rel_err(grg.matmul(x, direction), O.matmul(oracle_a, x.astype(np.float64), direction)) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("log_run", [0, 1, 2, 12])
@pytest.mark.parametrize("form", ["light", "merged", "tiled"])
def test_xcd_runs_light_path(art_a, oracle_a, dtype, log_run, form):
    """XCD runs of 2^log_run blocks (windows of 8 runs; blocks past the last full
    window, here all of them for 2^12, keep their order) on every per-row light path
    (segment by segment or merged) or tiled launch of both sweeps, with heavy and
    split rows beside it: every row produced once, exact for +-1 inputs, FP64 within
    1e-12, FP32c within 5e-7."""
    from pygrgl_spmv.backends.hip import HipRuntime, ffi
    from pygrgl_spmv.backends.hip.runtime import xcd_runs_flags

    layout = build_hip_layout([art_a], dtype=dtype, requirements=full_requirements(max_k_up=16, max_k_down=16))
    flags = xcd_runs_flags(log_run) | {"light": 0, "merged": ffi.FLAG_MERGED, "tiled": ffi.FLAG_TILED}[form]
    shape = {"unroll": 8, "rows_per_group": 1, "es": 2, "flags": flags}
    tol = FP64_TOL if dtype == np.float64 else FP32_TOL
    with HipRuntime(layout, fixed_shape=shape, split_min=40, split_chunk=16) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3390 + log_run)
        for k in (1, 16):
            for direction in ("up", "down"):
                x = rng.choice(np.array([-1.0, 1.0]), size=(k, _cols(oracle_a, direction))).astype(dtype)
                np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x.astype(np.float64), direction))
        y = rng.standard_normal((5, oracle_a.num_mutations))
        miss = rng.standard_normal((5, oracle_a.num_mutations))
        got = grg.matmul(y.astype(dtype), "down", miss=miss.astype(dtype))
        assert O.normwise_rel_err(got, O.matmul(oracle_a, y, "down", miss=miss)) <= tol


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("unroll", [4, 8, 16])
@pytest.mark.parametrize("es", [1, 4, 64])
def test_group_loaded_entries_match_per_lane_loads(art_a, oracle_a, dtype, unroll, es):
    """One row per lane group with the group's entry keys (merged down levels) or source
    indices (one-segment up levels and write-outs) loaded by the whole group and handed
    out by lane shuffles: bit for bit the sums of the per-lane loads (same order), exact
    against the oracle for +-1 inputs. es = 64 is one row per wave; k = 3 leaves lanes
    past the last column, which load entries but gather nothing; split and heavy rows
    beside it."""
    from pygrgl_spmv.backends.hip import HipRuntime, ffi

    layout = build_hip_layout([art_a], dtype=dtype, requirements=full_requirements(max_k_up=16, max_k_down=16))
    outs = {}
    for name, extra in (("per-lane", 0), ("group", ffi.FLAG_GROUP_KEYS)):
        for form, fl in (("merged", ffi.FLAG_MERGED), ("plain", 0)):
            shape = {"unroll": unroll, "rows_per_group": 1, "es": es, "flags": fl | extra}
            with HipRuntime(layout, fixed_shape=shape, split_min=40, split_chunk=16) as rt:
                (grg,) = rt.grgs
                r = np.random.default_rng(3460 + unroll + es)
                res = []
                for k in (1, 3, 16):
                    y = r.standard_normal((k, oracle_a.num_mutations)).astype(dtype)
                    m = r.standard_normal((k, oracle_a.num_mutations)).astype(dtype)
                    x = r.standard_normal((k, oracle_a.num_samples)).astype(dtype)
                    res.append(grg.matmul(y, "down", miss=m))
                    res.append(grg.matmul(y, "down", by_individual=True))
                    res.append(grg.matmul(x, "up"))
                for direction, cols in (("up", oracle_a.num_samples), ("down", oracle_a.num_mutations)):
                    t = r.choice(np.array([-1.0, 1.0]), size=(8, cols))
                    np.testing.assert_array_equal(grg.matmul(t.astype(dtype), direction), O.matmul(oracle_a, t, direction))
                outs[(name, form)] = res
    for form in ("merged", "plain"):
        for a, b in zip(outs[("per-lane", form)], outs[("group", form)]):
            np.testing.assert_array_equal(a, b)


def test_xcd_runs_flags_are_validated():
    from pygrgl_spmv.backends.hip.runtime import xcd_runs_flags

    assert xcd_runs_flags(4) == 128 | (4 << 8)
    with pytest.raises(ValueError, match="log_run"):
        xcd_runs_flags(16)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("order", ["plain", "remap", "runs"])
@pytest.mark.parametrize("kl,panel_rows,unroll", [(None, 256, 8), (1, 1024, 4), (2, 512, 8), (None, 1024, 4)])
def test_panel_launch(art_a, oracle_a, dtype, order, kl, panel_rows, unroll):
    """The panel form (1024-thread workgroups; rows of a panel share passes of
    LDS-staged sources) on every level of both sweeps, with heavy and split rows
    beside it (their own launch), many passes per panel (at most 7 staged sources per
    pass), 256-, 512- and 1024-row panels (1, 2 or 4 rows per lane group), two or
    three passes in flight, every block order, and one or several column chunks:
    exact for +-1 inputs, FP64 within 1e-12, FP32c within 5e-7."""
    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.backends.hip import ffi

    from pygrgl_spmv.backends.hip.runtime import xcd_runs_flags

    layout = build_hip_layout([art_a], dtype=dtype, requirements=full_requirements(max_k_up=16, max_k_down=16),
                              vram_budget_bytes=0)
    order_flags = {"plain": 0, "remap": ffi.FLAG_XCD_REMAP, "runs": xcd_runs_flags(1)}[order]
    shape = {"unroll": unroll, "rows_per_group": 1, "flags": ffi.FLAG_PANEL | order_flags}
    if kl is not None:
        shape["kl"] = kl
    tol = FP64_TOL if dtype == np.float64 else FP32_TOL
    with HipRuntime(layout, fixed_shape=shape, split_min=40, split_chunk=16, panels=True, panel_min_mean=0.0,
                    panel_min_rows=1, panel_rows=panel_rows, panel_staged=7) as rt:
        (grg,) = rt.grgs
        art = rt._artifacts[0]
        assert art.panels, "no level got panels"
        assert max(m[2] for m in art.panels.values()) > 3, "no panel with several passes"
        rng = np.random.default_rng(4410 + len(order) + (kl or 0) + panel_rows)
        for k in (2, 16):
            for direction in ("up", "down"):
                x = rng.choice(np.array([-1.0, 1.0]), size=(k, _cols(oracle_a, direction))).astype(dtype)
                np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x.astype(np.float64), direction))
        y = rng.standard_normal((5, oracle_a.num_mutations))
        miss = rng.standard_normal((5, oracle_a.num_mutations))
        got = grg.matmul(y.astype(dtype), "down", miss=miss.astype(dtype))
        assert O.normwise_rel_err(got, O.matmul(oracle_a, y, "down", miss=miss)) <= tol
        x = rng.standard_normal((8, oracle_a.num_individuals))
        got = grg.matmul(x.astype(dtype), "up", by_individual=True, init="xtx")
        assert O.normwise_rel_err(got, O.matmul(oracle_a, x, "up", by_individual=True, init="xtx")) <= tol


def test_panel_launch_tuned(art_a, oracle_a):
    """The tuner with panels on every level next to the other forms: whatever it picks
    matches the oracle. Panels need a layout without a VRAM budget."""
    from pygrgl_spmv.backends.hip import HipRuntime

    budgeted = build_hip_layout([art_a], requirements=full_requirements(max_k_up=16, max_k_down=16))
    with pytest.raises(ValueError, match="without a VRAM budget"):
        HipRuntime(budgeted, panels=True)

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=16, max_k_down=16), vram_budget_bytes=0)
    with HipRuntime(layout, panels=True, panel_min_mean=0.0, panel_min_rows=1) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(4420)
        for direction in ("up", "down"):
            x = rng.standard_normal((16, _cols(oracle_a, direction)))
            _check(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))


@pytest.mark.parametrize("unroll,rows", [(4, 1), (8, 1), (16, 1), (4, 2), (8, 2), (4, 4)])
@pytest.mark.parametrize("es", [1, 4])
@pytest.mark.parametrize("merged", [False, True])
def test_every_kernel_variant(art_a, oracle_a, unroll, rows, es, merged):
    """Each (gathers in flight, rows per lane group) variant of the light path, with the
    XCD remap, per segment or over merged entry keys, against the oracle in every
    sweep configuration."""
    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.backends.hip import ffi

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=8, max_k_down=8))
    shape = {"unroll": unroll, "rows_per_group": rows, "es": es,
             "flags": ffi.FLAG_XCD_REMAP | (ffi.FLAG_MERGED if merged else 0)}
    with HipRuntime(layout, fixed_shape=shape, split_min=40, split_chunk=16) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3350 + 10 * unroll + rows + es)
        for direction in ("up", "down"):
            x = rng.choice(np.array([-1.0, 1.0]), size=(5, _cols(oracle_a, direction)))
            np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))
        y = rng.standard_normal((3, oracle_a.num_mutations))
        miss = rng.standard_normal((3, oracle_a.num_mutations))
        _check(grg.matmul(y, "down", miss=miss), O.matmul(oracle_a, y, "down", miss=miss))
        x = rng.standard_normal((8, oracle_a.num_individuals))
        _check(grg.matmul(x, "up", by_individual=True, init="xtx"), O.matmul(oracle_a, x, "up", by_individual=True, init="xtx"))


@pytest.mark.parametrize("window,degree_window", [(8, 0), (64, 4)])
def test_min_child_node_order(art_a, oracle_a, window, degree_window):
    """Sweeps over the min-child device order (every level above 0 refined): exact for +-1
    inputs and within 1e-12 of the oracle in every sweep configuration, tuned."""
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=8, max_k_down=8))
    with HipRuntime(layout, cluster_window=window, cluster_degree_window=degree_window, cluster_min_rows=0,
                    degree_min_mean=0.0, split_min=40, split_chunk=16) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3470 + window)
        for direction in ("up", "down"):
            x = rng.choice(np.array([-1.0, 1.0]), size=(5, _cols(oracle_a, direction)))
            np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))
        y = rng.standard_normal((8, oracle_a.num_mutations))
        miss = rng.standard_normal((8, oracle_a.num_mutations))
        _check(grg.matmul(y, "down", miss=miss), O.matmul(oracle_a, y, "down", miss=miss))
        x = rng.standard_normal((8, oracle_a.num_individuals))
        _check(grg.matmul(x, "up", by_individual=True, init="xtx"), O.matmul(oracle_a, x, "up", by_individual=True, init="xtx"))
        x = rng.standard_normal((3, oracle_a.num_samples))
        _check(grg.matmul(x, "up", emit_all_nodes=True), O.matmul(oracle_a, x, "up", emit_all_nodes=True))


@pytest.mark.parametrize("unroll,rows", [(4, 1), (8, 1), (4, 2)])
def test_fp32_wide_lanes(art_a, oracle_a, unroll, rows):
    """FP32 values with 8 columns (32 bytes) per lane, k = 16 over 2 lanes."""
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], dtype=np.float32, requirements=full_requirements(max_k_up=16, max_k_down=16))
    shape = {"cpl": 8, "kl": 2, "es": 2, "unroll": unroll, "rows_per_group": rows}
    with HipRuntime(layout, fixed_shape=shape) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3370 + unroll + rows)
        for direction in ("up", "down"):
            x = rng.choice(np.array([-1.0, 1.0], dtype=np.float32), size=(16, _cols(oracle_a, direction)))
            np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x.astype(np.float64), direction))
            z = rng.standard_normal((16, _cols(oracle_a, direction))).astype(np.float32)
            assert O.normwise_rel_err(grg.matmul(z, direction), O.matmul(oracle_a, z.astype(np.float64), direction)) <= FP32_TOL


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("es,remap", [(1, False), (4, True), (16, False)])
@pytest.mark.parametrize("max_staged,max_rows", [(24, 16), (512, 1024)])
def test_tiled_level_launches(art_a, oracle_a, dtype, es, remap, max_staged, max_rows):
    """Every level launch in its LDS-staged tiled form (small tiles, direct tiles for rows
    longer than max_staged, heavy and split rows beside the tiles), against the oracle in
    every sweep configuration: exact for +-1 inputs, FP64 within 1e-12, FP32c within 5e-7."""
    from pygrgl_spmv.backends.hip import HipRuntime
    from pygrgl_spmv.backends.hip import ffi

    layout = build_hip_layout([art_a], dtype=dtype, requirements=full_requirements(max_k_up=16, max_k_down=16))
    assert layout.tiles and layout.bytes_by_category["tiles"] > 0
    flags = ffi.FLAG_TILED | (ffi.FLAG_XCD_REMAP if remap else 0)
    shape = {"es": es, "flags": flags}
    tol = FP64_TOL if dtype == np.float64 else FP32_TOL
    with HipRuntime(layout, fixed_shape=shape, split_min=40, split_chunk=16, tile_max_staged=max_staged,
                    tile_max_rows=max_rows) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3390 + es + max_staged)
        for k in (3, 16):
            for direction in ("up", "down"):
                x = rng.choice(np.array([-1.0, 1.0]), size=(k, _cols(oracle_a, direction))).astype(dtype)
                np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x.astype(np.float64), direction))
        y = rng.standard_normal((8, oracle_a.num_mutations))
        miss = rng.standard_normal((8, oracle_a.num_mutations))
        got = grg.matmul(y.astype(dtype), "down", miss=miss.astype(dtype))
        assert O.normwise_rel_err(got, O.matmul(oracle_a, y, "down", miss=miss)) <= tol
        x = rng.standard_normal((8, oracle_a.num_individuals))
        got = grg.matmul(x.astype(dtype), "up", by_individual=True, init="xtx")
        assert O.normwise_rel_err(got, O.matmul(oracle_a, x, "up", by_individual=True, init="xtx")) <= tol


def test_tuner_may_choose_tiles(art_a, oracle_a):
    """With tuning, each level picks the faster of its per-row and tiled forms; results
    agree with the oracle either way, and without tile tables no launch is tiled."""
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=16, max_k_down=16))
    rng = np.random.default_rng(3399)
    x = rng.standard_normal((16, oracle_a.num_samples))
    y = rng.standard_normal((16, oracle_a.num_mutations))
    for tiles in (True, False):
        with HipRuntime(layout, tiles=tiles) as rt:
            (grg,) = rt.grgs
            _check(grg.matmul(x, "up"), O.matmul(oracle_a, x, "up"))
            _check(grg.matmul(y, "down"), O.matmul(oracle_a, y, "down"))
            assert tiles or not rt._artifacts[0].tiles


# ---------------------------------------------------------------------------
# row-group launches (FP64, MFMA)
# ---------------------------------------------------------------------------


def _group_runtime(layout, shape, **kw):
    from pygrgl_spmv.backends.hip import HipRuntime

    # groups on every level launch with segments, whatever their sharing
    return HipRuntime(layout, fixed_shape=shape, group_min_mean=0.0, group_min_rows=1, group_max_fill=100.0, **kw)


@pytest.mark.parametrize("cpl,unroll", [(1, 4), (2, 8), (4, 16)])
@pytest.mark.parametrize("order", ["plain", "remap", "runs"])
def test_row_group_launches(art_a, oracle_a, cpl, unroll, order):
    """Every level launch of both sweeps as a row-group launch (16-row groups summed by FP64
    MFMA, 16-column chunks), heavy and split rows beside it, in each XCD block order:
    exact for +-1 inputs, within 1e-12 for N(0, 1), with missing data, by-individual
    columns and the init modes; k from 1 to 33 covers partial and several column chunks."""
    from pygrgl_spmv.backends.hip import ffi
    from pygrgl_spmv.backends.hip.runtime import xcd_runs_flags

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=36, max_k_down=36))
    extra = {"plain": 0, "remap": ffi.FLAG_XCD_REMAP, "runs": xcd_runs_flags(2)}[order]
    shape = {"unroll": unroll, "rows_per_group": 1, "es": 1, "cpl": cpl, "kl": 16 // cpl, "flags": ffi.FLAG_GROUPS | extra}
    with _group_runtime(layout, shape, split_min=40, split_chunk=16) as rt:
        (grg,) = rt.grgs
        assert rt._artifacts[0].groups, "row groups must be built"
        rng = np.random.default_rng(3950 + cpl + unroll)
        for k in (1, 4, 16, 20, 36) if cpl == 4 else (1, 2, 16, 33):
            if k % cpl:
                continue
            for direction in ("up", "down"):
                x = rng.choice(np.array([-1.0, 1.0]), size=(k, _cols(oracle_a, direction)))
                np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))
        y = rng.standard_normal((4, oracle_a.num_mutations))
        miss = rng.standard_normal((4, oracle_a.num_mutations))
        _check(grg.matmul(y, "down", miss=miss.copy()), O.matmul(oracle_a, y, "down", miss=miss.copy()))
        x = rng.standard_normal((4, oracle_a.num_individuals))
        _check(grg.matmul(x, "up", by_individual=True, init="xtx"), O.matmul(oracle_a, x, "up", by_individual=True, init="xtx"))
        y = rng.standard_normal((4, oracle_a.num_mutations))
        init = rng.standard_normal((4, oracle_a.num_nodes))
        _check(grg.matmul(y, "down", by_individual=True, init=init), O.matmul(oracle_a, y, "down", by_individual=True, init=init))


def test_row_groups_cover_levels_and_tuner_may_choose_them(art_a, oracle_a):
    """Groups are built for the FP64 level launches (not for FP32c runtimes), every group
    entry set and store mask is consistent with the level's rows, and a tuned runtime (the
    row-group form among its candidates) matches the oracle."""
    from pygrgl_spmv.backends.hip import HipRuntime

    for dtype in (np.float64, np.float32):
        layout = build_hip_layout([art_a], dtype=dtype, requirements=full_requirements(max_k_up=16, max_k_down=16))
        with HipRuntime(layout, group_min_mean=0.0, group_min_rows=1, group_max_fill=100.0) as rt:
            art = rt._artifacts[0]
            assert bool(art.groups) == (dtype == np.float64)
            for (cfg, level), (n_groups, n_ent) in art.groups.items():
                rows = int(art.level_offsets[level + 1] - art.level_offsets[level])
                assert n_groups == -(-rows // 16) and n_ent % 4 == 0
            (grg,) = rt.grgs
            rng = np.random.default_rng(3960)
            x = rng.standard_normal((16, oracle_a.num_samples)).astype(dtype)
            y = rng.standard_normal((16, oracle_a.num_mutations)).astype(dtype)
            tol = FP64_TOL if dtype == np.float64 else FP32_TOL
            assert O.normwise_rel_err(grg.matmul(x, "up"), O.matmul(oracle_a, x.astype(np.float64), "up")) <= tol
            assert O.normwise_rel_err(grg.matmul(y, "down"), O.matmul(oracle_a, y.astype(np.float64), "down")) <= tol


# ---------------------------------------------------------------------------
# streamed blocks through the pinned ring
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("ring", [1, 2, 4])
def test_all_streamed_ring(art_a, oracle_a, ring):
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=4, max_k_down=4),
                              ring_buffer_size=ring, allow_residency=False)
    assert not any(b.resident for a in layout.artifacts for b in (*a.blocks_up, *a.blocks_down))
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3400 + ring)
        for direction in ("up", "down"):
            x = rng.standard_normal((4, _cols(oracle_a, direction)))
            _check(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))
        y = rng.standard_normal((2, oracle_a.num_mutations))
        miss = rng.standard_normal((2, oracle_a.num_mutations))
        _check(grg.matmul(y, "down", miss=miss), O.matmul(oracle_a, y, "down", miss=miss))
        x = rng.standard_normal((2, oracle_a.num_samples))
        _check(grg.matmul(x, "up", init="xtx"), O.matmul(oracle_a, x, "up", init="xtx"))


def test_partially_resident_budget(art_a, oracle_a):
    """A budget between the smallest streamed layout and full residency keeps some
    blocks resident and streams the rest (planner rules as triton/backend.py:439-499)."""
    from pygrgl_spmv.backends.hip import HipRuntime

    req = full_requirements(max_k_up=2, max_k_down=2)
    full = build_hip_layout([art_a], requirements=req, vram_budget_bytes=0)
    streamed = build_hip_layout([art_a], requirements=req, ring_buffer_size=2, allow_residency=False)
    budget = (streamed.bytes_total + full.required_budget_for_full_residency) // 2
    layout = build_hip_layout([art_a], requirements=req, vram_budget_bytes=budget, ring_buffer_size=2)
    flags = [b.resident for a in layout.artifacts for b in (*a.blocks_up, *a.blocks_down)]
    assert any(flags) and not all(flags)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        rng = np.random.default_rng(3500)
        for direction in ("up", "down"):
            x = rng.choice(np.array([-1.0, 1.0]), size=(2, _cols(oracle_a, direction)))
            np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))


def test_streamed_band_closed_form(artifact_dir):
    from pygrgl_spmv.backends.hip import HipRuntime

    n, bw = 64, 8
    art = write_band_artifact(artifact_dir, "band-n64", n=n, bandwidth=bw)
    layout = build_hip_layout([art], requirements=full_requirements(max_k_up=1, max_k_down=1),
                              ring_buffer_size=2, allow_residency=False)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        x = np.random.default_rng(3600).choice(np.array([-1.0, 1.0]), size=(1, n))
        np.testing.assert_array_equal(grg.matmul(x, "up"), O.band_expected_up(x.T, shifts=(0, bw, 2 * bw), bandwidth=bw).T)
        np.testing.assert_array_equal(grg.matmul(x, "down"), O.band_expected_down(x.T, shifts=(0, bw, 2 * bw), bandwidth=bw).T)


def _overlaps(compute, copies) -> bool:
    """True when some timed slot copy ran while a timed streamed launch was in flight."""
    for c0, c1 in compute:
        for h0, h1 in copies:
            if h0.elapsed_time(c1) > 0.0 and c0.elapsed_time(h1) > 0.0:
                return True
    return False


@pytest.mark.parametrize("ring", [1, 2])
def test_ring_copies_overlap_compute(art_a, oracle_a, monkeypatch, ring):
    """Slot copies run on their own streams: with one slot each copy waits for the launch
    that read the slot before, so no copy overlaps a launch; with two slots the copy of
    the next block runs under the current launch. Each streamed launch is followed by a
    device spin so that a launch window is long enough to see; the results stay exact.
    Counterpart: ``runtime/test_triton.py:363-441`` (timed copy and launch wrappers)."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=2, max_k_down=2),
                              ring_buffer_size=ring, allow_residency=False)
    assert layout.allocated_ring_buffer_size == ring
    state = {"active": True, "compute": [], "copy": []}
    launch, copy = HipRuntime._launch_streamed_block, HipRuntime._copy_block_to_slot

    def timed_launch(self, args, stream_ptr):
        if not state["active"]:
            return launch(self, args, stream_ptr)
        root = self._root_stream
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record(root)
        launch(self, args, stream_ptr)
        with torch.cuda.stream(root):
            torch.cuda._sleep(5_000_000)
        t1.record(root)
        state["compute"].append((t0, t1))

    def timed_copy(self, step, prev_compute_done, copy_done):
        if not state["active"]:
            return copy(self, step, prev_compute_done, copy_done)
        cs = self._slot_copy_streams[step.slot]
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        if prev_compute_done is not None:
            cs.wait_event(prev_compute_done)
        t0.record(cs)
        copy(self, step, None, copy_done)
        t1.record(cs)
        state["copy"].append((t0, t1))

    monkeypatch.setattr(HipRuntime, "_launch_streamed_block", timed_launch)
    monkeypatch.setattr(HipRuntime, "_copy_block_to_slot", timed_copy)
    rng = np.random.default_rng(3700 + ring)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        for i, direction in enumerate(("up", "down", "up")):
            x = rng.choice(np.array([-1.0, 1.0]), size=(2, _cols(oracle_a, direction)))
            np.testing.assert_array_equal(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))
            torch.cuda.synchronize()
            if i == 0:
                assert len(state["compute"]) >= 3 and len(state["copy"]) == len(state["compute"])
                assert _overlaps(state["compute"], state["copy"]) == (ring > 1)
                state["active"] = False


# ---------------------------------------------------------------------------
# hipGraph capture of prepared op chains
# ---------------------------------------------------------------------------


def _chain_expected(g, order, src, init_first, init_second):
    mid = O.matmul(g, src, order[0], init=init_first)
    return O.matmul(g, mid, order[1], init=init_second)


def _enter_chain(stack, grg, order, k):
    first = stack.enter_context(grg.prepare_matmul_cuda(direction=order[0], k=k, init_mode="vector"))
    second = stack.enter_context(grg.prepare_matmul_cuda(direction=order[1], k=k, init_mode="vector"))
    assert first.output.data_ptr() == second.input.data_ptr()
    return first, second


def _run_chain(first, second, src, i1, i2, result):
    first.input.copy_(src)
    first.init_vector.copy_(i1)
    first()
    second.init_vector.copy_(i2)
    second()
    result.copy_(second.output)


def test_streamed_ops_are_not_capturable(art_a):
    """Sweeps with streamed blocks run eagerly only; resident sweeps are capturable."""
    from pygrgl_spmv.backends.hip import HipRuntime

    req = full_requirements(max_k_up=1, max_k_down=1)
    for streamed in (False, True):
        layout = build_hip_layout([art_a], requirements=req, ring_buffer_size=2 if streamed else 0, allow_residency=not streamed)
        with HipRuntime(layout) as rt:
            (grg,) = rt.grgs
            for direction in ("up", "down"):
                assert grg.prepare_matmul_cuda(direction=direction, k=1).capturable is (not streamed)


@pytest.mark.skip(reason="capturing a sweep with streamed blocks ended in a host segmentation fault inside "
                         "hipStreamEndCapture (pinned-host slot copies on per-slot streams forked from the launch "
                         "stream); no cause has been shown in this code, so the runtime refuses such a capture and "
                         "this case stays off the GPU")
def test_graph_capture_streamed_chain(art_a, oracle_a):
    """Counterpart of the reference's streamed down -> up capture (``test_triton.py:945-990``):
    every block streamed through a 2-slot ring, the chain captured once and replayed against the
    oracle."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    with torch.cuda.device(0):
        cap = torch.cuda.Stream()
    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=1, max_k_down=1), stream=cap,
                              ring_buffer_size=2, allow_residency=False)
    order = ("down", "up")
    src_np = np.arange(_cols(oracle_a, order[0]), dtype=np.float64).reshape(1, -1)
    i1_np, i2_np = np.asarray([0.5]), np.asarray([1.5])
    expected = _chain_expected(oracle_a, order, src_np, i1_np, i2_np)
    with HipRuntime(layout) as rt, ExitStack() as stack:
        (grg,) = rt.grgs
        first, second = _enter_chain(stack, grg, order, 1)
        dev = first.input.device
        src = torch.from_numpy(src_np).to(dev)
        i1, i2 = torch.from_numpy(i1_np).to(dev), torch.from_numpy(i2_np).to(dev)
        result = torch.empty_like(second.output)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cap):
            _run_chain(first, second, src, i1, i2, result)
        for rep in range(2):
            graph.replay()
            torch.cuda.synchronize()
            _check(result.cpu().numpy(), expected)


def test_graph_capture_chain_replays(art_a, oracle_a):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    with torch.cuda.device(0):
        cap = torch.cuda.Stream()
    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=2, max_k_down=2), stream=cap)
    order = ("down", "up")
    rng = np.random.default_rng(3700)
    src_np = rng.standard_normal((2, _cols(oracle_a, order[0])))
    i1_np, i2_np = rng.standard_normal(2), rng.standard_normal(2)
    expected = _chain_expected(oracle_a, order, src_np, i1_np, i2_np)
    with HipRuntime(layout) as rt, ExitStack() as stack:
        (grg,) = rt.grgs
        first, second = _enter_chain(stack, grg, order, 2)
        dev = first.input.device
        src = torch.from_numpy(src_np).to(dev)
        i1, i2 = torch.from_numpy(i1_np).to(dev), torch.from_numpy(i2_np).to(dev)
        result = torch.empty_like(second.output)
        # eager reference run on the capture stream
        with torch.cuda.stream(cap):
            _run_chain(first, second, src, i1, i2, result)
        torch.cuda.synchronize()
        eager = result.cpu().numpy().copy()
        _check(eager, expected)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cap):
            _run_chain(first, second, src, i1, i2, result)
        for rep in range(3):
            # unrelated eager work on the arena between replays
            first.input.zero_()
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(result.cpu().numpy(), eager, err_msg=f"replay {rep}")


def test_graph_capture_multi_grg_mixed_heights(artifact_dir, art_b):
    """Two GRGs of different heights (4-level band, synthetic) share one runtime and
    one captured graph (reference test_triton.py:978-1014)."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    band = write_band_artifact(artifact_dir, "band-h4-n16", n=16, bandwidth=4, levels=4)
    arts = [band, art_b]
    oracles = [O.load_artifact(p) for p in arts]
    with torch.cuda.device(0):
        cap = torch.cuda.Stream()
    layout = build_hip_layout(arts, requirements=full_requirements(max_k_up=1, max_k_down=1), stream=cap)
    orders = [("up", "down"), ("down", "up")]
    cases = []
    for i, (g, order) in enumerate(zip(oracles, orders)):
        src = np.arange(_cols(g, order[0]), dtype=np.float64).reshape(1, -1) + 100.0 * i
        i1, i2 = np.array([0.25 + i]), np.array([0.75 + i])
        cases.append((order, src, i1, i2, _chain_expected(g, order, src, i1, i2)))
    with HipRuntime(layout) as rt, ExitStack() as stack:
        chains = [_enter_chain(stack, grg, case[0], 1) for grg, case in zip(rt.grgs, cases)]
        dev = chains[0][0].input.device
        tens = [(torch.from_numpy(c[1]).to(dev), torch.from_numpy(c[2]).to(dev), torch.from_numpy(c[3]).to(dev)) for c in cases]
        results = [torch.empty_like(ch[1].output) for ch in chains]
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cap):
            for (first, second), (src, i1, i2), res in zip(chains, tens, results):
                _run_chain(first, second, src, i1, i2, res)
        for rep in range(2):
            graph.replay()
            torch.cuda.synchronize()
            for case, res in zip(cases, results):
                _check(res.cpu().numpy(), case[4])


# ---------------------------------------------------------------------------
# several GRGs, mixed k, concurrent runtimes, lifecycle
# ---------------------------------------------------------------------------


def test_multi_grg_mixed_k(artifact_dir, art_a, art_b):
    from pygrgl_spmv.backends.hip import HipRuntime

    band = write_band_artifact(artifact_dir, "band-n64", n=64, bandwidth=8)
    arts = [art_a, art_b, band]
    oracles = [O.load_artifact(p) for p in arts]
    layout = build_hip_layout(arts, requirements=full_requirements(max_k_up=16, max_k_down=16))
    with HipRuntime(layout) as rt:
        assert len(rt.grgs) == 3
        rng = np.random.default_rng(3800)
        for grg, g, k in zip(rt.grgs, oracles, (1, 7, 16)):
            for direction in ("up", "down"):
                x = rng.standard_normal((k, _cols(g, direction)))
                _check(grg.matmul(x, direction), O.matmul(g, x, direction))


@pytest.mark.parametrize("persistent", [True, False])
def test_batched_sweep_of_several_grgs(artifact_dir, art_a, art_b, persistent):
    """Up and down sweeps of three GRGs with mixed k (missing data, by-individual,
    xtx init among them) from one job table: one persistent launch with grid-wide
    barriers, or one launch per phase; every output against the oracle, and a
    capped persistent grid (workgroups looping over each phase) gives the same."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    band = write_band_artifact(artifact_dir, "band-n64-batch", n=64, bandwidth=8)
    arts = [art_a, art_b, band]
    oracles = [O.load_artifact(p) for p in arts]
    layout = build_hip_layout(arts, requirements=full_requirements(max_k_up=16, max_k_down=16))
    rng = np.random.default_rng(3850)
    with HipRuntime(layout) as rt:
        ga, gb, gc = rt.grgs
        oa, ob, oc = oracles
        entries = [
            dict(grg=ga, direction="up", k=5), dict(grg=ga, direction="down", k=16, use_miss=True),
            dict(grg=gb, direction="up", k=1), dict(grg=gb, direction="down", k=7),
            dict(grg=gc, direction="up", k=3), dict(grg=gc, direction="down", k=2),
            dict(grg=ga, direction="up", k=4, by_individual=True, init_mode="xtx"),
        ]
        grids = [None, 7] if persistent else [None]
        for grid in grids:
            batch = rt.prepare_batched_sweep(entries, persistent=persistent, grid=grid)
            assert batch.num_phases >= max(g.height for g in oracles)
            inputs = []
            for e, view, g in zip(entries, batch.inputs, (oa, oa, ob, ob, oc, oc, oa)):
                cols = g.num_individuals if e.get("by_individual") else _cols(g, e["direction"])
                x = rng.standard_normal((e["k"], cols))
                view.copy_(torch.from_numpy(x))
                inputs.append(x)
            miss = rng.standard_normal((16, oa.num_mutations))
            batch.miss_inputs[1].copy_(torch.from_numpy(miss))
            batch()
            batch.check()
            want = [O.matmul(oa, inputs[0], "up"), O.matmul(oa, inputs[1], "down", miss=miss),
                    O.matmul(ob, inputs[2], "up"), O.matmul(ob, inputs[3], "down"),
                    O.matmul(oc, inputs[4], "up"), O.matmul(oc, inputs[5], "down"),
                    O.matmul(oa, inputs[6], "up", by_individual=True, init="xtx")]
            for out, w in zip(batch.outputs, want):
                _check(out.cpu().numpy(), w)


@pytest.mark.parametrize("ring,residency", [(2, False), (3, True)])
def test_batched_sweep_with_streamed_blocks(artifact_dir, art_a, art_b, ring, residency):
    """The grouped batched form on layouts that stream blocks through the pinned ring
    (every block streamed through 2 slots, or a budget that keeps part of them
    resident with 3 slots): up and down sweeps of three GRGs with mixed k, the
    streamed blocks copied and accumulated between phases, every output against the
    oracle on two consecutive runs; the persistent form is refused and the batch is
    not capturable."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    band = write_band_artifact(artifact_dir, "band-n64-batch-streamed", n=64, bandwidth=8)
    arts = [art_a, art_b, band]
    oracles = [O.load_artifact(p) for p in arts]
    req = full_requirements(max_k_up=16, max_k_down=16)
    full = build_hip_layout(arts, requirements=req, vram_budget_bytes=0)
    budget = 0
    if residency:
        sparse = int(full.bytes_by_category.get("resident_sparse", 0))
        budget = int(full.required_budget_for_full_residency) - sparse // 2
    layout = build_hip_layout(arts, requirements=req, ring_buffer_size=ring, allow_residency=residency,
                              vram_budget_bytes=budget)
    n_streamed = sum(not b.resident for a in layout.artifacts for b in (*a.blocks_up, *a.blocks_down))
    assert n_streamed > 0
    rng = np.random.default_rng(3870 + ring)
    with HipRuntime(layout) as rt:
        ga, gb, gc = rt.grgs
        oa, ob, oc = oracles
        entries = [
            dict(grg=ga, direction="up", k=5), dict(grg=ga, direction="down", k=16, use_miss=True),
            dict(grg=gb, direction="up", k=1), dict(grg=gb, direction="down", k=7),
            dict(grg=gc, direction="up", k=3), dict(grg=gc, direction="down", k=2),
        ]
        with pytest.raises(ValueError, match="persistent batched sweep needs every block"):
            rt.prepare_batched_sweep(entries, persistent=True)
        batch = rt.prepare_batched_sweep(entries, persistent=False)
        assert not batch.capturable
        for _ in range(2):
            inputs = []
            for e, view, g in zip(entries, batch.inputs, (oa, oa, ob, ob, oc, oc)):
                x = rng.standard_normal((e["k"], _cols(g, e["direction"])))
                view.copy_(torch.from_numpy(x))
                inputs.append(x)
            miss = rng.standard_normal((16, oa.num_mutations))
            batch.miss_inputs[1].copy_(torch.from_numpy(miss))
            batch()
            batch.check()
            want = [O.matmul(oa, inputs[0], "up"), O.matmul(oa, inputs[1], "down", miss=miss),
                    O.matmul(ob, inputs[2], "up"), O.matmul(ob, inputs[3], "down"),
                    O.matmul(oc, inputs[4], "up"), O.matmul(oc, inputs[5], "down")]
            for out, w in zip(batch.outputs, want):
                _check(out.cpu().numpy(), w)


def test_persistent_full_grid_and_barrier_cost(artifact_dir, art_a, art_b):
    """The persistent launch at every co-resident workgroup (the grid where each XCD
    holds most members per sub-group): every output against the oracle on three
    consecutive runs, and the fixed costs of the two-tier barrier stay small (about
    9 us per barrier and 15 us per launch at 1024 workgroups on MI355X; the bound
    leaves room for a noisy box but fails the one-release-per-workgroup form, which
    cost about 85 us)."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    arts = [art_a, art_b]
    oracles = [O.load_artifact(p) for p in arts]
    layout = build_hip_layout(arts, requirements=full_requirements(max_k_up=16, max_k_down=16))
    rng = np.random.default_rng(3855)
    with HipRuntime(layout) as rt:
        ga, gb = rt.grgs
        entries = [dict(grg=ga, direction="up", k=9), dict(grg=ga, direction="down", k=16, use_miss=True),
                   dict(grg=gb, direction="up", k=2), dict(grg=gb, direction="down", k=5)]
        probe = rt.prepare_batched_sweep(entries, persistent=True)
        batch = rt.prepare_batched_sweep(entries, persistent=True, grid=probe.max_grid)
        assert batch.grid == batch.max_grid
        gs = (oracles[0], oracles[0], oracles[1], oracles[1])
        for rep in range(3):
            xs = []
            for e, view, g in zip(entries, batch.inputs, gs):
                x = rng.choice(np.array([-1.0, 1.0]), size=(e["k"], _cols(g, e["direction"])))
                view.copy_(torch.from_numpy(x))
                xs.append(x)
            miss = rng.choice(np.array([-1.0, 1.0]), size=(16, oracles[0].num_mutations))
            batch.miss_inputs[1].copy_(torch.from_numpy(miss))
            batch()
            batch.check()
            for i, (e, out, x, g) in enumerate(zip(entries, batch.outputs, xs, gs)):
                want = O.matmul(g, x, e["direction"], miss=miss if i == 1 else None)
                np.testing.assert_array_equal(out.cpu().numpy(), want)
        costs = batch.barrier_costs()
        print(f"grid {batch.grid}: {costs}")
        assert costs["barrier_us"] <= 25.0 and costs["one_phase_us"] <= 100.0


def test_tuned_shapes_follow_leading_dimensions(art_a, oracle_a):
    """A batched sweep at k = 4 (private buffers with ld = 4, where 4 columns per lane are
    admissible) is tuned first; an op of the same k on the runtime's arena (ld = max_k = 6,
    so at most 2 columns per lane) must get a shape of its own, not the cached one."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=6, max_k_down=6))
    rng = np.random.default_rng(3860)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        for direction in ("up", "down"):
            batch = rt.prepare_batched_sweep([dict(grg=grg, direction=direction, k=4)], persistent=False)
            x = rng.standard_normal((4, _cols(oracle_a, direction)))
            batch.inputs[0].copy_(torch.from_numpy(x))
            batch()
            batch.check()
            _check(batch.outputs[0].cpu().numpy(), O.matmul(oracle_a, x, direction))
            _check(grg.matmul(x, direction), O.matmul(oracle_a, x, direction))


def test_tuned_shapes_reused_from_disk(tmp_path, art_a, oracle_a):
    """The first runtime tunes and writes its shapes; a second runtime over the same artifact
    takes every level and write-out shape from the disk cache, with the same results."""
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=8, max_k_down=8))
    rng = np.random.default_rng(3870)
    x = rng.standard_normal((8, oracle_a.num_samples))
    y = rng.standard_normal((8, oracle_a.num_mutations))
    logs = []
    for _ in range(2):
        with HipRuntime(layout, tune_cache=tmp_path) as rt:
            (grg,) = rt.grgs
            _check(grg.matmul(x, "up"), O.matmul(oracle_a, x, "up"))
            _check(grg.matmul(y, "down"), O.matmul(oracle_a, y, "down"))
            logs.append(dict(rt.tune_log))
    assert list(tmp_path.glob("*.json"))
    assert logs[0] and not any(v.get("cached") for v in logs[0].values())
    assert set(logs[1]) == set(logs[0]) and all(v.get("cached") for v in logs[1].values())
    assert all(logs[1][key]["best"][1:] == logs[0][key]["best"][1:] for key in logs[0])


def test_batched_sweep_captured_and_replayed(art_a, art_b):
    """A persistent batched up + down pair captured into one hipGraph: every replay
    (with eager work and synchronizations between replays) equals the eager run
    bit for bit, and +-1 inputs give the oracle's integers exactly."""
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    arts = [art_a, art_b]
    oracles = [O.load_artifact(p) for p in arts]
    with torch.cuda.device(0):
        stream = torch.cuda.Stream()
    layout = build_hip_layout(arts, requirements=full_requirements(max_k_up=8, max_k_down=8), stream=stream)
    rng = np.random.default_rng(3860)
    with HipRuntime(layout) as rt:
        ga, gb = rt.grgs
        up = rt.prepare_batched_sweep([dict(grg=ga, direction="up", k=8), dict(grg=gb, direction="up", k=3)])
        dn = rt.prepare_batched_sweep([dict(grg=ga, direction="down", k=2), dict(grg=gb, direction="down", k=8)])
        xs = [rng.choice(np.array([-1.0, 1.0]), size=(8, oracles[0].num_samples)),
              rng.choice(np.array([-1.0, 1.0]), size=(3, oracles[1].num_samples))]
        ys = [rng.choice(np.array([-1.0, 1.0]), size=(2, oracles[0].num_mutations)),
              rng.choice(np.array([-1.0, 1.0]), size=(8, oracles[1].num_mutations))]
        dev_x = [torch.from_numpy(x).to(rt.device) for x in xs]
        dev_y = [torch.from_numpy(y).to(rt.device) for y in ys]

        def run():
            for v, x in zip(up.inputs, dev_x):
                v.copy_(x)
            for v, y in zip(dn.inputs, dev_y):
                v.copy_(y)
            up()
            dn()

        with torch.cuda.stream(stream):
            run()
        torch.cuda.synchronize()
        eager = [o.cpu().clone() for o in (*up.outputs, *dn.outputs)]
        for o, g, x in zip(up.outputs, oracles, xs):
            np.testing.assert_array_equal(o.cpu().numpy(), O.matmul(g, x, "up"))
        for o, g, y in zip(dn.outputs, oracles, ys):
            np.testing.assert_array_equal(o.cpu().numpy(), O.matmul(g, y, "down"))
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            run()
        for rep in range(3):
            for o in (*up.outputs, *dn.outputs):
                o.zero_()
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            up.check()
            got = [o.cpu() for o in (*up.outputs, *dn.outputs)]
            assert all(torch.equal(a, b) for a, b in zip(got, eager)), f"replay {rep} differs from the eager run"


def test_two_runtimes_on_separate_streams(art_a, art_b):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime

    with torch.cuda.device(0):
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    req = full_requirements(max_k_up=3, max_k_down=3)
    oa, ob = O.load_artifact(art_a), O.load_artifact(art_b)
    with HipRuntime(build_hip_layout([art_a], requirements=req, stream=s1)) as r1, \
            HipRuntime(build_hip_layout([art_b], requirements=req, stream=s2)) as r2:
        (ga,), (gb,) = r1.grgs, r2.grgs
        rng = np.random.default_rng(3900)
        xa = rng.standard_normal((3, oa.num_samples))
        xb = rng.standard_normal((3, ob.num_samples))
        with ExitStack() as stack:
            pa = stack.enter_context(ga.prepare_matmul_cuda(direction="up", k=3))
            pb = stack.enter_context(gb.prepare_matmul_cuda(direction="up", k=3))
            with torch.cuda.stream(s1):
                pa.input.copy_(torch.from_numpy(xa).to(pa.input.device))
            with torch.cuda.stream(s2):
                pb.input.copy_(torch.from_numpy(xb).to(pb.input.device))
            # one call each: an op's output shares the arena with its input
            pa()
            pb()
            torch.cuda.synchronize()
            _check(pa.output.cpu().numpy(), O.matmul(oa, xa, "up"))
            _check(pb.output.cpu().numpy(), O.matmul(ob, xb, "up"))


def test_runtime_lifecycle(art_a):
    from pygrgl_spmv.backends.hip import HipRuntime

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=2, max_k_down=2))
    rt = HipRuntime(layout)
    with pytest.raises(RuntimeError):
        _ = rt.grgs
    with rt:
        (grg,) = rt.grgs
        with rt._call_scope():
            with pytest.raises(RuntimeError, match="concurrent runtime.grgs calls are not supported"):
                with rt._call_scope():
                    pass
        up = grg.prepare_matmul_cuda(direction="up", k=2)
        down = grg.prepare_matmul_cuda(direction="down", k=2)
        assert up.output.data_ptr() == down.input.data_ptr()
        assert down.output.data_ptr() == up.input.data_ptr()
    with pytest.raises(RuntimeError):
        _ = rt.grgs


def test_disabled_direction_rejected(art_a):
    from pygrgl_spmv.backends.hip import HipRuntime
    from tests._builders import hip_pair

    layout = build_hip_layout([art_a], requirements=full_requirements(max_k_up=2, max_k_down=2), pair=hip_pair(down=False))
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        with pytest.raises(ValueError, match="DOWN plan is not configured"):
            grg.prepare_matmul_cuda(direction="down", k=1)


def test_chromosome_batch_captured_in_one_graph(artifact_dir):
    """Several GRGs with mixed k in one runtime: one captured graph for every op,
    replays identical to the eager run, smallest GRGs against the oracle."""
    import torch  # noqa: F401

    from pygrgl_spmv.chromosomes import mixed_k, run
    from pygrgl_spmv.synth import chromosome_workloads

    cfgs = chromosome_workloads(5, 400, seed=77)
    cfgs = [c.__class__(**{**c.__dict__, "num_mutations": max(500, c.num_mutations // 200), "num_trees": 30}) for c in cfgs]
    arts = [write_synthetic_artifact(artifact_dir, f"chrom-{i}", c) for i, c in enumerate(cfgs)]
    res = run(arts, budget_bytes=2_000_000_000, ring=4, replays=3)
    assert res["captured"] and res["graph_replays_checked"] == 3
    assert res["ks"] == [mixed_k(i) for i in range(5)]
