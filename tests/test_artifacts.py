"""``.grg_spmv`` format contract (mirrors the reference's ``runtime/test_artifacts.py``)."""

from __future__ import annotations

from pathlib import Path
import warnings
import zipfile

import numpy as np
import pytest
import scipy.sparse as sp

import pygrgl_spmv.grg.artifact as artifact_module
from pygrgl_spmv.backends.base import split_selector_by_level
from pygrgl_spmv.grg.artifact import (
    GRG_SPMV_FORMAT_MAGIC,
    GRG_SPMV_FORMAT_VERSION,
    iter_artifact_blocks,
    load_grg_spmv,
    save_grg_spmv,
    scan_grg_spmv,
)
from tests._builders import write_band_artifact, write_synthetic_artifact


@pytest.fixture(scope="module")
def synth_artifact(artifact_dir):
    from pygrgl_spmv.synth import SynthConfig

    cfg = SynthConfig(num_samples=300, num_mutations=2000, seed=3, ploidy=2, num_trees=100, keep_fraction=0.3,
                      window_fraction=0.1, num_segments=2, missing_fraction=0.05, missing_pool=4)
    return write_synthetic_artifact(artifact_dir, "art-synth", cfg)


def test_split_selector_by_level_toy():
    """Reference test_artifacts.py:25-52 (unsorted selector, empty level)."""
    offsets = np.asarray([0, 3, 3, 7, 10], dtype=np.int64)
    rows = np.asarray([2, 0, 2, 1, 2, 0])
    cols = np.asarray([8, 1, 4, 6, 2, 9])
    sel = sp.csr_matrix((np.ones(rows.size, dtype=np.int8), (rows, cols)), shape=(4, 10))
    pairs = split_selector_by_level(sel, offsets)
    assert pairs[1][0].size == 0 and pairs[1][1].size == 0
    rr = np.concatenate([r for r, _ in pairs])
    cc = np.concatenate([c.astype(np.int64) + offsets[l] for l, (_, c) in enumerate(pairs)])
    rebuilt = sp.csr_matrix((np.ones(sel.nnz, dtype=np.int8), (rr, cc)), shape=sel.shape)
    np.testing.assert_array_equal(rebuilt.toarray(), sel.toarray())


def test_split_selector_wide_column_dtype():
    wide = int(np.iinfo(np.int32).max) + 1
    sel = sp.csr_matrix((np.ones(1, dtype=np.int8), (np.asarray([0]), np.asarray([wide]))), shape=(1, wide + 1))
    ((rows, cols),) = split_selector_by_level(sel, np.asarray([0, wide + 1], dtype=np.int64))
    assert rows.dtype == np.int32 and cols.dtype == np.int64
    np.testing.assert_array_equal(cols, [wide])


def test_scan_matches_loaded_state(synth_artifact):
    scan = scan_grg_spmv(synth_artifact)
    state = load_grg_spmv(synth_artifact, np.float64)
    assert scan.path == Path(synth_artifact).resolve()
    for f in ("num_samples", "num_mutations", "num_nodes", "num_individuals", "num_edges", "ploidy", "has_missing_data"):
        assert getattr(scan, f) == getattr(state, f)
    assert scan.has_xtx_bias == (state.init_xtx_up_bias is not None)
    np.testing.assert_array_equal(scan.level_offsets, state.level_offsets)
    assert scan.num_levels == len(state.level_offsets) - 1
    assert scan.has_individual_coals == (state.coalescence_counts is not None)
    assert scan.selector_nnz_by_level.shape == (scan.num_levels, 2)
    assert int(scan.selector_nnz_by_level[:, 0].sum()) == scan.selector_mut_nnz == state.sel_mut.nnz
    assert int(scan.selector_nnz_by_level[:, 1].sum()) == scan.selector_miss_nnz == state.sel_miss.nnz
    for level, (rows, _) in enumerate(split_selector_by_level(state.sel_mut, state.level_offsets)):
        assert int(scan.selector_nnz_by_level[level, 0]) == rows.size
    loaded = {(d, s): b for d, row in enumerate(state.A_blocks) for s, b in enumerate(row)}
    assert len(scan.blocks) == len(loaded)
    for b in scan.blocks:
        ref = loaded[(b.dst_level, b.src_level)]
        assert b.shape == ref.shape and b.nnz == ref.nnz
        assert b.indices_dtype == ref.indices.dtype and b.indptr_dtype == ref.indptr.dtype
    assert sum(b.nnz for b in scan.blocks) == scan.num_edges


def test_iter_blocks_match_loaded(synth_artifact):
    state = load_grg_spmv(synth_artifact, np.float64)
    loaded = {(d, s): b for d, row in enumerate(state.A_blocks) for s, b in enumerate(row)}
    blocks = list(iter_artifact_blocks(synth_artifact))
    assert len(blocks) == len(loaded)
    for b in blocks:
        np.testing.assert_array_equal(b.indices, loaded[(b.dst_level, b.src_level)].indices)
        np.testing.assert_array_equal(b.indptr, loaded[(b.dst_level, b.src_level)].indptr)


def test_loaded_blocks_share_readonly_data(synth_artifact):
    state = load_grg_spmv(synth_artifact, np.float64)
    data = [b.data for row in state.A_blocks for b in row if b.nnz > 0]
    assert data and data[0].strides == (0,) and not data[0].flags.writeable


def test_save_is_atomic_when_savez_fails(tmp_path, monkeypatch, synth_artifact):
    state = load_grg_spmv(synth_artifact, np.float64)
    target = tmp_path / "atomic.grg_spmv"

    def boom(*args, **kwargs):
        raise RuntimeError("savez exploded")

    monkeypatch.setattr(artifact_module.np, "savez", boom)
    with pytest.raises(RuntimeError, match="savez exploded"):
        save_grg_spmv(state, target)
    assert not target.exists()
    assert list(tmp_path.iterdir()) == []


def test_archive_is_uncompressed_with_direct_scan_keys(synth_artifact):
    with zipfile.ZipFile(synth_artifact) as zf:
        infos = zf.infolist()
        names = {i.filename for i in infos}
    assert all(i.compress_type == zipfile.ZIP_STORED for i in infos)
    for key in ("scan_block_levels", "scan_block_shapes", "scan_block_nnzs", "scan_block_struct_itemsize",
                "scan_selector_nnzs", "scan_selector_nnz_by_level", "scan_metadata_struct_itemsize"):
        assert f"{key}.npy" in names
    assert "sel_mut_nnz_by_level.npy" not in names


def test_scan_does_not_load_structural_arrays(tmp_path, monkeypatch):
    """Reference test_artifacts.py:231-252: n=4, bw=1 band -> blocks [(1,0,4),(2,0,4),(2,1,4)]."""
    art = write_band_artifact(tmp_path, "direct-scan", n=4, bandwidth=1)
    real = artifact_module._read_struct

    def guarded(npz, key):
        if key != "level_offsets":
            raise AssertionError(f"scan loaded structural array {key}")
        return real(npz, key)

    artifact_module.clear_scan_cache()
    monkeypatch.setattr(artifact_module, "_read_struct", guarded)
    scan = scan_grg_spmv(art)
    assert scan.num_levels == 3
    assert scan.selector_mut_nnz == 4 and scan.selector_miss_nnz == 0
    assert [(b.dst_level, b.src_level, b.nnz) for b in scan.blocks] == [(1, 0, 4), (2, 0, 4), (2, 1, 4)]


@pytest.mark.parametrize("version", [5, 6, 7])
def test_old_versions_rejected(tmp_path, version):
    art = tmp_path / "old.grg_spmv"
    with art.open("wb") as fh:
        np.savez(fh, grg_spmv_magic=np.asarray(GRG_SPMV_FORMAT_MAGIC), grg_spmv_format_version=np.asarray(version, dtype=np.int32))
    with pytest.raises(ValueError, match=f"got version {version}, expected {GRG_SPMV_FORMAT_VERSION}"):
        scan_grg_spmv(art)


def test_bad_magic_rejected(tmp_path):
    art = tmp_path / "bad.grg_spmv"
    with art.open("wb") as fh:
        np.savez(fh, grg_spmv_magic=np.asarray("nope"), grg_spmv_format_version=np.asarray(8, dtype=np.int32))
    with pytest.raises(ValueError, match="got magic"):
        scan_grg_spmv(art)


def test_wide_storage_warns_on_load(tmp_path):
    art = write_band_artifact(tmp_path, "wide-storage", n=4, bandwidth=1, struct_dtype=np.int64)
    with pytest.warns(RuntimeWarning, match="wider than necessary"):
        state = load_grg_spmv(art, np.float64)
    assert all(b.indices.dtype == np.int32 for row in state.A_blocks for b in row)


def test_generated_artifact_invariants(synth_artifact):
    """Compile invariants (reference compile.py:56-90,183-188,370-382) hold for generated GRGs."""
    state = load_grg_spmv(synth_artifact, np.float64)
    S = state.num_samples
    np.testing.assert_array_equal(np.asarray(state.node_perm[:S]), np.arange(S))
    assert int(state.level_offsets[1]) >= S
    for d, row in enumerate(state.A_blocks):
        for s, b in enumerate(row):
            assert s < d
            for r in range(b.shape[0]):
                cols = b.indices[b.indptr[r] : b.indptr[r + 1]]
                assert np.all(np.diff(cols) > 0), "columns must be sorted and unique"
    # every internal node has a child one level down (levels are longest paths)
    off = np.asarray(state.level_offsets)
    for d in range(1, len(off) - 1):
        reach = np.zeros(off[d + 1] - off[d], dtype=bool)
        reach |= np.diff(state.A_blocks[d][d - 1].indptr) > 0
        assert reach.all()


def test_written_keys_and_dtypes_follow_format_8(synth_artifact):
    """Every key of format 8 (reference artifact.py:124-184) with its storage dtype; nothing else."""
    with np.load(synth_artifact, allow_pickle=False) as z:
        files = set(z.files)
        got = {k: z[k].dtype for k in files}
        levels = z["scan_block_levels"]
    counts = ("num_samples", "num_mutations", "num_nodes", "ploidy", "num_individuals", "num_edges")
    flags = ("has_missing_data", "has_individual_coals", "has_xtx_bias")
    structs = ("level_offsets", "node_perm", "sample_to_individual", "sel_mut_indices", "sel_mut_indptr",
               "sel_miss_indices", "sel_miss_indptr")
    fixed = {"grg_spmv_format_version": np.int32, "mutation_positions": np.float64, "mutation_times": np.float64,
             "coalescence_counts": np.int64, "scan_block_levels": np.int32, "scan_block_shapes": np.int64,
             "scan_block_nnzs": np.int64, "scan_block_struct_itemsize": np.uint8, "scan_selector_nnzs": np.int64,
             "scan_selector_nnz_by_level": np.int64, "scan_metadata_struct_itemsize": np.uint8}
    free = ("grg_spmv_magic", "mutation_alleles", "mutation_allele_offsets", "mutation_ref_alleles",
            "mutation_ref_allele_offsets", "init_vector_up_bias", "init_vector_down_bias", "init_xtx_up_bias",
            "init_xtx_down_bias")
    blocks = {f"A_blocks_{d}_{s}_{part}" for d, s in levels.tolist() for part in ("indices", "indptr")}
    want = set(counts) | set(flags) | set(structs) | set(fixed) | set(free) | blocks
    assert files == want
    for k in counts:
        assert got[k] == np.int64
    for k in flags:
        assert got[k] == np.bool_
    for k in structs + tuple(blocks):
        assert got[k] in (np.int32, np.int64)
    for k, dt in fixed.items():
        assert got[k] == dt, k
    # a block for every level pair d > s, empty ones included
    state = load_grg_spmv(synth_artifact, np.float64)
    L = len(state.level_offsets) - 1
    assert sorted(map(tuple, levels.tolist())) == [(d, s) for d in range(L) for s in range(d)]


def test_presence_flags_follow_optional_keys(tmp_path, synth_artifact):
    """Dropping the coalescence counts and the XTX biases clears both flags, and the loader
    then returns None for them."""
    state = load_grg_spmv(synth_artifact, np.float64)
    assert state.coalescence_counts is not None and state.init_xtx_up_bias is not None
    state.coalescence_counts = None
    state.init_xtx_down_bias = None
    out = tmp_path / "no-optional.grg_spmv"
    save_grg_spmv(state, out)
    with np.load(out, allow_pickle=False) as z:
        assert not bool(z["has_individual_coals"]) and not bool(z["has_xtx_bias"])
        assert "coalescence_counts" not in z.files and "init_xtx_down_bias" not in z.files
    back = load_grg_spmv(out, np.float64)
    assert back.coalescence_counts is None and back.init_xtx_up_bias is None and back.init_xtx_down_bias is None
    scan = scan_grg_spmv(out)
    assert not scan.has_individual_coals and not scan.has_xtx_bias
