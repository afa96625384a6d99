"""``.grg_spmv`` artifact persistence, metadata scan and streamed block loading.

Format version 8, an uncompressed ``np.savez`` archive with the key set of the
reference's ``pygrgl_spmv/grg/artifact.py:124-184`` (format keys, counts, flags,
structure, mutation table, init biases, selectors, one ``A_blocks_{d}_{s}``
CSR pair per level pair including empty ones, and ``scan_*`` metadata so a
planner can size everything without reading structural arrays). The key names
and dtypes must match the reference byte-for-byte so artifacts are
interchangeable between backends.
"""

from __future__ import annotations

from dataclasses import dataclass
import functools
import logging
import os
from pathlib import Path
import tempfile
import time
import warnings

import numpy as np

from pygrgl_spmv.grg.compile import CompiledOperatorState, _invert_permutation
from pygrgl_spmv.grg.sparse import binary_csr_from_parts

GRG_SPMV_FORMAT_MAGIC = "grg_spmv"
GRG_SPMV_FORMAT_VERSION = 8
_MAGIC_KEY = "grg_spmv_magic"
_VERSION_KEY = "grg_spmv_format_version"
_LOGGER = logging.getLogger(__name__)


def artifact_path_for_grg(grg_path: Path, artifact_root: Path) -> Path:
    """Artifact path mirroring the resolved source path under ``artifact_root``."""
    resolved = Path(grg_path).expanduser().resolve()
    parts = ["_abs", *resolved.parts[1:]] if resolved.is_absolute() else ["_rel", *resolved.parts]
    return Path(artifact_root).joinpath(*parts).with_suffix(".grg_spmv")


@dataclass(frozen=True)
class ArtifactBlockScan:
    dst_level: int
    src_level: int
    shape: tuple[int, int]
    nnz: int
    indices_dtype: np.dtype
    indptr_dtype: np.dtype


@dataclass(frozen=True)
class ArtifactBlock:
    dst_level: int
    src_level: int
    shape: tuple[int, int]
    indices: np.ndarray
    indptr: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.indices.size)


@dataclass(frozen=True)
class ArtifactScan:
    path: Path
    num_samples: int
    num_mutations: int
    num_nodes: int
    num_individuals: int
    num_edges: int
    ploidy: int
    has_missing_data: bool
    has_xtx_bias: bool
    level_offsets: np.ndarray
    level_sizes: tuple[int, ...]
    num_levels: int
    selector_mut_nnz: int
    selector_miss_nnz: int
    has_individual_coals: bool
    selector_nnz_by_level: np.ndarray
    node_perm_dtype: np.dtype
    sample_to_individual_dtype: np.dtype
    blocks: tuple[ArtifactBlockScan, ...]

    @property
    def max_block_nnz(self) -> int:
        return max((b.nnz for b in self.blocks), default=0)

    @property
    def sel_mut_nnz_by_level(self) -> np.ndarray:
        return self.selector_nnz_by_level[:, 0]

    @property
    def sel_miss_nnz_by_level(self) -> np.ndarray:
        return self.selector_nnz_by_level[:, 1]


def _selector_nnz_by_level(indices: np.ndarray, level_offsets: np.ndarray) -> np.ndarray:
    offsets = np.asarray(level_offsets)
    n_levels = max(int(offsets.size) - 1, 0)
    idx = np.asarray(indices)
    if idx.size == 0:
        return np.zeros(n_levels, dtype=np.int64)
    lv = np.searchsorted(offsets, idx, side="right") - 1
    if np.any(lv < 0) or np.any(lv >= n_levels):
        raise ValueError("selector indices are outside level_offsets")
    return np.bincount(lv, minlength=n_levels).astype(np.int64)


def save_grg_spmv(state: CompiledOperatorState, artifact_path) -> None:
    """Write ``state`` atomically (temp file + fsync + rename), as ``artifact.py:109-204``."""
    if state.A_blocks is None:
        raise RuntimeError("compiled operator blocks are required for .grg_spmv save")
    if state.init_vector_up_bias is None or state.init_vector_down_bias is None:
        raise RuntimeError("init vector bias cache must be populated before .grg_spmv save")
    t0 = time.perf_counter()
    offsets = np.asarray(state.level_offsets)
    sel_mut_idx = np.asarray(state.sel_mut.indices)
    sel_miss_idx = np.asarray(state.sel_miss.indices)
    by_level = np.column_stack(
        (_selector_nnz_by_level(sel_mut_idx, offsets), _selector_nnz_by_level(sel_miss_idx, offsets))
    ).astype(np.int64)
    node_perm = np.asarray(state.node_perm)
    s2i = np.asarray(state.sample_to_individual)
    arrays: dict[str, np.ndarray] = {
        _MAGIC_KEY: np.asarray(GRG_SPMV_FORMAT_MAGIC),
        _VERSION_KEY: np.asarray(GRG_SPMV_FORMAT_VERSION, dtype=np.int32),
        "num_samples": np.asarray(state.num_samples, dtype=np.int64),
        "num_mutations": np.asarray(state.num_mutations, dtype=np.int64),
        "num_nodes": np.asarray(state.num_nodes, dtype=np.int64),
        "ploidy": np.asarray(state.ploidy, dtype=np.int64),
        "num_individuals": np.asarray(state.num_individuals, dtype=np.int64),
        "num_edges": np.asarray(state.num_edges, dtype=np.int64),
        "has_missing_data": np.asarray(state.has_missing_data, dtype=bool),
        "level_offsets": offsets,
        "node_perm": node_perm,
        "sample_to_individual": s2i,
        "mutation_positions": np.asarray(state.mutation_positions, dtype=np.float64),
        "mutation_times": np.asarray(state.mutation_times, dtype=np.float64),
        "mutation_alleles": np.asarray(state.mutation_alleles),
        "mutation_allele_offsets": np.asarray(state.mutation_allele_offsets),
        "mutation_ref_alleles": np.asarray(state.mutation_ref_alleles),
        "mutation_ref_allele_offsets": np.asarray(state.mutation_ref_allele_offsets),
        "has_individual_coals": np.asarray(state.coalescence_counts is not None, dtype=bool),
        "init_vector_up_bias": np.asarray(state.init_vector_up_bias),
        "init_vector_down_bias": np.asarray(state.init_vector_down_bias),
        "has_xtx_bias": np.asarray(
            state.init_xtx_up_bias is not None and state.init_xtx_down_bias is not None, dtype=bool
        ),
        "sel_mut_indices": sel_mut_idx,
        "sel_mut_indptr": np.asarray(state.sel_mut.indptr),
        "sel_miss_indices": sel_miss_idx,
        "sel_miss_indptr": np.asarray(state.sel_miss.indptr),
    }
    if state.coalescence_counts is not None:
        arrays["coalescence_counts"] = np.asarray(state.coalescence_counts, dtype=np.int64)
    if state.init_xtx_up_bias is not None:
        arrays["init_xtx_up_bias"] = np.asarray(state.init_xtx_up_bias)
    if state.init_xtx_down_bias is not None:
        arrays["init_xtx_down_bias"] = np.asarray(state.init_xtx_down_bias)
    levels, shapes, nnzs, itemsizes = [], [], [], []
    for d, row in enumerate(state.A_blocks):
        for s, block in enumerate(row):
            idx = np.asarray(block.indices)
            ptr = np.asarray(block.indptr)
            arrays[f"A_blocks_{d}_{s}_indices"] = idx
            arrays[f"A_blocks_{d}_{s}_indptr"] = ptr
            levels.append((d, s))
            shapes.append((int(block.shape[0]), int(block.shape[1])))
            nnzs.append(int(idx.size))
            itemsizes.append((idx.dtype.itemsize, ptr.dtype.itemsize))
    arrays["scan_block_levels"] = np.asarray(levels, dtype=np.int32).reshape(-1, 2)
    arrays["scan_block_shapes"] = np.asarray(shapes, dtype=np.int64).reshape(-1, 2)
    arrays["scan_block_nnzs"] = np.asarray(nnzs, dtype=np.int64)
    arrays["scan_block_struct_itemsize"] = np.asarray(itemsizes, dtype=np.uint8).reshape(-1, 2)
    arrays["scan_selector_nnzs"] = np.asarray((state.sel_mut.nnz, state.sel_miss.nnz), dtype=np.int64)
    arrays["scan_selector_nnz_by_level"] = by_level
    arrays["scan_metadata_struct_itemsize"] = np.asarray(
        (node_perm.dtype.itemsize, s2i.dtype.itemsize), dtype=np.uint8
    )
    path = Path(artifact_path)
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp_name = tempfile.mkstemp(prefix=f".{path.name}.", suffix=".tmp", dir=path.parent)
    tmp = Path(tmp_name)
    try:
        with os.fdopen(fd, "wb") as handle:
            np.savez(handle, **arrays)
            handle.flush()
            os.fsync(handle.fileno())
        tmp.replace(path)
    except BaseException:
        tmp.unlink(missing_ok=True)
        raise
    _scan_grg_spmv_cached.cache_clear()
    _LOGGER.debug("save_grg_spmv path=%s blocks=%d elapsed=%.3fs", path, len(levels), time.perf_counter() - t0)


def _open_archive(artifact_path):
    return np.load(artifact_path, allow_pickle=False)


def _validate_archive(data, artifact_path) -> None:
    if _MAGIC_KEY not in data.files:
        raise ValueError(f"unsupported .grg_spmv artifact in {artifact_path}: missing {_MAGIC_KEY}")
    magic = str(np.asarray(data[_MAGIC_KEY]).item())
    if magic != GRG_SPMV_FORMAT_MAGIC:
        raise ValueError(f"unsupported .grg_spmv artifact in {artifact_path}: got magic {magic!r}")
    if _VERSION_KEY not in data.files:
        raise ValueError(f"unsupported .grg_spmv artifact in {artifact_path}: missing {_VERSION_KEY}")
    version = int(np.asarray(data[_VERSION_KEY]).item())
    if version != GRG_SPMV_FORMAT_VERSION:
        raise ValueError(
            f"unsupported .grg_spmv artifact in {artifact_path}: got version {version}, expected {GRG_SPMV_FORMAT_VERSION}"
        )


def _load_struct_array(data, key: str, *, non_negative: bool = True) -> np.ndarray:
    arr = np.asarray(data[key])
    if arr.dtype not in (np.dtype(np.int32), np.dtype(np.int64)):
        raise ValueError(f"unsupported structural dtype for {key}: {arr.dtype}")
    if non_negative and arr.size and int(arr.min()) < 0:
        raise ValueError(f"structural array {key} must be non-negative")
    return arr


def _struct_dtype_from_itemsize(itemsize: int) -> np.dtype:
    if int(itemsize) == 4:
        return np.dtype(np.int32)
    if int(itemsize) == 8:
        return np.dtype(np.int64)
    raise ValueError(f"unsupported structural itemsize: {itemsize}")


def _scalar(data, key: str):
    return np.asarray(data[key]).item()


@functools.lru_cache(maxsize=None)
def _scan_grg_spmv_cached(artifact_path: Path) -> ArtifactScan:
    with _open_archive(artifact_path) as data:
        _validate_archive(data, artifact_path)
        offsets = _load_struct_array(data, "level_offsets")
        offsets.setflags(write=False)
        sizes = tuple(int(offsets[i + 1]) - int(offsets[i]) for i in range(len(offsets) - 1))
        b_levels = np.asarray(data["scan_block_levels"])
        b_shapes = np.asarray(data["scan_block_shapes"])
        b_nnzs = np.asarray(data["scan_block_nnzs"])
        b_items = np.asarray(data["scan_block_struct_itemsize"])
        sel_nnzs = np.asarray(data["scan_selector_nnzs"])
        sel_by_level = np.asarray(data["scan_selector_nnz_by_level"], dtype=np.int64)
        meta_items = np.asarray(data["scan_metadata_struct_itemsize"])
        for name, arr in (("scan_block_levels", b_levels), ("scan_block_shapes", b_shapes), ("scan_block_struct_itemsize", b_items)):
            if arr.ndim != 2 or arr.shape[1] != 2:
                raise ValueError(f"{name} must have shape (num_blocks, 2), got {arr.shape}")
        if b_nnzs.ndim != 1:
            raise ValueError(f"scan_block_nnzs must be one-dimensional, got {b_nnzs.shape}")
        nb = int(b_levels.shape[0])
        if b_shapes.shape[0] != nb or b_nnzs.shape[0] != nb or b_items.shape[0] != nb:
            raise ValueError("scan block metadata arrays must have matching lengths")
        if sel_nnzs.shape != (2,):
            raise ValueError(f"scan_selector_nnzs must have shape (2,), got {sel_nnzs.shape}")
        if sel_by_level.shape != (len(sizes), 2):
            raise ValueError(f"scan_selector_nnz_by_level must have shape ({len(sizes)}, 2), got {sel_by_level.shape}")
        if meta_items.shape != (2,):
            raise ValueError(f"scan_metadata_struct_itemsize must have shape (2,), got {meta_items.shape}")
        if any(arr.size and np.any(arr < 0) for arr in (b_levels, b_shapes, b_nnzs, sel_nnzs, sel_by_level)):
            raise ValueError("scan metadata arrays must be non-negative")
        if np.any(sel_by_level.sum(axis=0) != sel_nnzs):
            raise ValueError("scan selector metadata arrays must have matching nnz totals")
        sel_by_level.setflags(write=False)
        blocks = tuple(
            ArtifactBlockScan(
                dst_level=int(lv[0]),
                src_level=int(lv[1]),
                shape=(int(sh[0]), int(sh[1])),
                nnz=int(nz),
                indices_dtype=_struct_dtype_from_itemsize(int(it[0])),
                indptr_dtype=_struct_dtype_from_itemsize(int(it[1])),
            )
            for lv, sh, nz, it in zip(b_levels, b_shapes, b_nnzs, b_items)
        )
        return ArtifactScan(
            path=artifact_path,
            num_samples=int(_scalar(data, "num_samples")),
            num_mutations=int(_scalar(data, "num_mutations")),
            num_nodes=int(_scalar(data, "num_nodes")),
            num_individuals=int(_scalar(data, "num_individuals")),
            num_edges=int(_scalar(data, "num_edges")),
            ploidy=int(_scalar(data, "ploidy")),
            has_missing_data=bool(_scalar(data, "has_missing_data")),
            has_xtx_bias=bool(_scalar(data, "has_xtx_bias")),
            level_offsets=offsets,
            level_sizes=sizes,
            num_levels=len(sizes),
            selector_mut_nnz=int(sel_nnzs[0]),
            selector_miss_nnz=int(sel_nnzs[1]),
            has_individual_coals=bool(_scalar(data, "has_individual_coals")),
            selector_nnz_by_level=sel_by_level,
            node_perm_dtype=_struct_dtype_from_itemsize(int(meta_items[0])),
            sample_to_individual_dtype=_struct_dtype_from_itemsize(int(meta_items[1])),
            blocks=blocks,
        )


def scan_grg_spmv(path) -> ArtifactScan:
    """Metadata-only scan (cached per resolved path; cleared by ``save_grg_spmv``)."""
    return _scan_grg_spmv_cached(Path(path).resolve())


def iter_artifact_blocks(path):
    """Yield every stored level block, one at a time, in stored order."""
    artifact_path = Path(path)
    with _open_archive(artifact_path) as data:
        _validate_archive(data, artifact_path)
        b_levels = np.asarray(data["scan_block_levels"])
        b_shapes = np.asarray(data["scan_block_shapes"])
        if b_levels.ndim != 2 or b_levels.shape[1] != 2:
            raise ValueError(f"scan_block_levels must have shape (num_blocks, 2), got {b_levels.shape}")
        if b_shapes.ndim != 2 or b_shapes.shape[1] != 2:
            raise ValueError(f"scan_block_shapes must have shape (num_blocks, 2), got {b_shapes.shape}")
        if b_levels.shape[0] != b_shapes.shape[0]:
            raise ValueError("scan block metadata arrays must have matching lengths")
        for (d, s), shape in zip(b_levels, b_shapes):
            d, s = int(d), int(s)
            yield ArtifactBlock(
                dst_level=d,
                src_level=s,
                shape=(int(shape[0]), int(shape[1])),
                indices=_load_struct_array(data, f"A_blocks_{d}_{s}_indices"),
                indptr=_load_struct_array(data, f"A_blocks_{d}_{s}_indptr"),
            )


def _load_grg_spmv_host(artifact_path, dtype) -> CompiledOperatorState:
    """Everything except the level blocks (``A_blocks=None``)."""
    with _open_archive(artifact_path) as data:
        _validate_archive(data, artifact_path)
        S = int(_scalar(data, "num_samples"))
        M = int(_scalar(data, "num_mutations"))
        N = int(_scalar(data, "num_nodes"))
        node_perm = _load_struct_array(data, "node_perm")
        coal = None
        if bool(_scalar(data, "has_individual_coals")):
            coal = np.asarray(data["coalescence_counts"], dtype=np.int64)
        xtx_up = xtx_down = None
        if bool(_scalar(data, "has_xtx_bias")):
            xtx_up = np.asarray(data["init_xtx_up_bias"], dtype=dtype)
            xtx_down = np.asarray(data["init_xtx_down_bias"], dtype=dtype)
        return CompiledOperatorState(
            A_blocks=None,
            level_offsets=_load_struct_array(data, "level_offsets"),
            node_perm=node_perm,
            inv_node_perm=_invert_permutation(node_perm),
            sel_mut=binary_csr_from_parts(
                indices=_load_struct_array(data, "sel_mut_indices"),
                indptr=_load_struct_array(data, "sel_mut_indptr"),
                shape=(M, N),
            ),
            sel_miss=binary_csr_from_parts(
                indices=_load_struct_array(data, "sel_miss_indices"),
                indptr=_load_struct_array(data, "sel_miss_indptr"),
                shape=(M, N),
            ),
            num_samples=S,
            num_mutations=M,
            num_nodes=N,
            ploidy=int(_scalar(data, "ploidy")),
            num_individuals=int(_scalar(data, "num_individuals")),
            num_edges=int(_scalar(data, "num_edges")),
            has_missing_data=bool(_scalar(data, "has_missing_data")),
            sample_to_individual=_load_struct_array(data, "sample_to_individual"),
            mutation_positions=np.asarray(data["mutation_positions"], dtype=np.float64),
            mutation_times=np.asarray(data["mutation_times"], dtype=np.float64),
            mutation_alleles=np.asarray(data["mutation_alleles"]),
            mutation_allele_offsets=np.asarray(data["mutation_allele_offsets"]),
            mutation_ref_alleles=np.asarray(data["mutation_ref_alleles"]),
            mutation_ref_allele_offsets=np.asarray(data["mutation_ref_allele_offsets"]),
            coalescence_counts=coal,
            init_vector_up_bias=np.asarray(data["init_vector_up_bias"], dtype=dtype),
            init_vector_down_bias=np.asarray(data["init_vector_down_bias"], dtype=dtype),
            init_xtx_up_bias=xtx_up,
            init_xtx_down_bias=xtx_down,
        )


def load_grg_spmv(artifact_path, dtype) -> CompiledOperatorState:
    """Host state plus every level block as a binary CSR (narrowed dtypes)."""
    header = _load_grg_spmv_host(artifact_path, dtype)
    n_levels = len(header.level_offsets) - 1
    sizes = np.diff(np.asarray(header.level_offsets, dtype=np.int64))
    grid: list[list] = [[None] * d for d in range(n_levels)]
    wasted = False
    for block in iter_artifact_blocks(artifact_path):
        loaded = binary_csr_from_parts(indices=block.indices, indptr=block.indptr, shape=block.shape, shared_data=True)
        if (
            loaded.indices.dtype.itemsize < block.indices.dtype.itemsize
            or loaded.indptr.dtype.itemsize < block.indptr.dtype.itemsize
        ):
            wasted = True
        grid[block.dst_level][block.src_level] = loaded
    for d in range(n_levels):
        for s in range(d):
            if grid[d][s] is None:
                grid[d][s] = binary_csr_from_parts(
                    indices=np.empty(0, dtype=np.int32),
                    indptr=np.zeros(int(sizes[d]) + 1, dtype=np.int32),
                    shape=(int(sizes[d]), int(sizes[s])),
                )
    header.A_blocks = grid
    if wasted:
        warnings.warn(
            f"{artifact_path} stores some structural arrays wider than necessary; "
            "load_grg_spmv() rebuilt smaller in-memory dtypes, so the artifact uses more disk space than needed",
            RuntimeWarning,
            stacklevel=2,
        )
    return header


__all__ = [
    "ArtifactBlock",
    "ArtifactBlockScan",
    "ArtifactScan",
    "GRG_SPMV_FORMAT_MAGIC",
    "GRG_SPMV_FORMAT_VERSION",
    "artifact_path_for_grg",
    "iter_artifact_blocks",
    "load_grg_spmv",
    "save_grg_spmv",
    "scan_grg_spmv",
]
