"""``.grg_spmv`` artifacts: one schema table drives the writer, the scan and the loader.

Format version 8 is an uncompressed ``np.savez`` archive. Its key names and
dtypes are the reference's (``pygrgl_spmv/grg/artifact.py:124-184``), so the
same artifact loads in either backend. Everything else here is organised
around ``_HEADER``, a table of the per-artifact keys: each row names the key,
how its value is stored and where it lives in ``CompiledOperatorState``. The
writer walks the table to build the archive, the loader walks it to rebuild
the state, and both check a key's presence rule from the same row. The level
blocks (``A_blocks_{d}_{s}_indices`` / ``_indptr``) and the planner metadata
(``scan_*``, readable without touching a structural array) have tables of
their own (``_block_keys``, ``_SCAN_SHAPES``).
"""

from __future__ import annotations

from contextlib import contextmanager
from dataclasses import dataclass
import logging
import os
from pathlib import Path
import time
import warnings

import numpy as np

from pygrgl_spmv.grg.compile import CompiledOperatorState, _invert_permutation
from pygrgl_spmv.grg.sparse import binary_csr_from_parts

GRG_SPMV_FORMAT_MAGIC = "grg_spmv"
GRG_SPMV_FORMAT_VERSION = 8
_LOGGER = logging.getLogger(__name__)

# ---------------------------------------------------------------------------
# The schema
# ---------------------------------------------------------------------------

# How a key's value is stored.
_COUNT = "count"      # int64 scalar
_FLAG = "flag"        # bool scalar
_STRUCT = "struct"    # non-negative int32 or int64 array (narrowest that holds the values)
_F64 = "f64"          # float64 array
_AS_IS = "as_is"      # stored with the writer's dtype, read back unchanged
_VALUES = "values"    # float array, stored as written, read in the caller's value dtype
_I64 = "i64"          # int64 array


@dataclass(frozen=True)
class _Key:
    """One archive key: ``kind`` is its storage, ``attr`` the state attribute it holds
    (a ``sel_*`` CSR part is ``"sel_mut.indices"``), ``when`` the state attribute whose
    presence (not None) decides whether the key is written (None: always)."""

    name: str
    kind: str
    attr: str
    when: str | None = None


_HEADER: tuple[_Key, ...] = (
    _Key("num_samples", _COUNT, "num_samples"),
    _Key("num_mutations", _COUNT, "num_mutations"),
    _Key("num_nodes", _COUNT, "num_nodes"),
    _Key("ploidy", _COUNT, "ploidy"),
    _Key("num_individuals", _COUNT, "num_individuals"),
    _Key("num_edges", _COUNT, "num_edges"),
    _Key("has_missing_data", _FLAG, "has_missing_data"),
    _Key("level_offsets", _STRUCT, "level_offsets"),
    _Key("node_perm", _STRUCT, "node_perm"),
    _Key("sample_to_individual", _STRUCT, "sample_to_individual"),
    _Key("mutation_positions", _F64, "mutation_positions"),
    _Key("mutation_times", _F64, "mutation_times"),
    _Key("mutation_alleles", _AS_IS, "mutation_alleles"),
    _Key("mutation_allele_offsets", _AS_IS, "mutation_allele_offsets"),
    _Key("mutation_ref_alleles", _AS_IS, "mutation_ref_alleles"),
    _Key("mutation_ref_allele_offsets", _AS_IS, "mutation_ref_allele_offsets"),
    _Key("init_vector_up_bias", _VALUES, "init_vector_up_bias"),
    _Key("init_vector_down_bias", _VALUES, "init_vector_down_bias"),
    _Key("sel_mut_indices", _STRUCT, "sel_mut.indices"),
    _Key("sel_mut_indptr", _STRUCT, "sel_mut.indptr"),
    _Key("sel_miss_indices", _STRUCT, "sel_miss.indices"),
    _Key("sel_miss_indptr", _STRUCT, "sel_miss.indptr"),
    _Key("coalescence_counts", _I64, "coalescence_counts", when="coalescence_counts"),
    _Key("init_xtx_up_bias", _VALUES, "init_xtx_up_bias", when="init_xtx_up_bias"),
    _Key("init_xtx_down_bias", _VALUES, "init_xtx_down_bias", when="init_xtx_down_bias"),
)

# Presence flags of the optional keys: flag -> the keys that must all be present.
_PRESENCE_FLAGS = {
    "has_individual_coals": ("coalescence_counts",),
    "has_xtx_bias": ("init_xtx_up_bias", "init_xtx_down_bias"),
}

# Planner metadata: key -> (storage dtype, shape with "blocks" / "levels" as named extents).
_SCAN_SHAPES = {
    "scan_block_levels": (np.int32, ("blocks", 2)),
    "scan_block_shapes": (np.int64, ("blocks", 2)),
    "scan_block_nnzs": (np.int64, ("blocks",)),
    "scan_block_struct_itemsize": (np.uint8, ("blocks", 2)),
    "scan_selector_nnzs": (np.int64, (2,)),
    "scan_selector_nnz_by_level": (np.int64, ("levels", 2)),
    "scan_metadata_struct_itemsize": (np.uint8, (2,)),
}

_STRUCT_DTYPES = {4: np.dtype(np.int32), 8: np.dtype(np.int64)}


def _block_keys(dst: int, src: int) -> tuple[str, str]:
    """Archive keys of level block A[dst][src]: (indices, indptr)."""
    return f"A_blocks_{dst}_{src}_indices", f"A_blocks_{dst}_{src}_indptr"


def _struct_dtype(itemsize: int) -> np.dtype:
    try:
        return _STRUCT_DTYPES[int(itemsize)]
    except KeyError:
        raise ValueError(f"unsupported structural itemsize: {itemsize}") from None


def _state_value(state: CompiledOperatorState, attr: str):
    obj = state
    for part in attr.split("."):
        obj = getattr(obj, part)
    return obj


def _encode(kind: str, value) -> np.ndarray:
    if kind == _COUNT:
        return np.asarray(int(value), dtype=np.int64)
    if kind == _FLAG:
        return np.asarray(bool(value), dtype=bool)
    if kind == _F64:
        return np.asarray(value, dtype=np.float64)
    if kind == _I64:
        return np.asarray(value, dtype=np.int64)
    return np.asarray(value)  # structural arrays keep the compiler's narrowed dtype


# ---------------------------------------------------------------------------
# Paths and scan records
# ---------------------------------------------------------------------------


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


# ---------------------------------------------------------------------------
# Writing
# ---------------------------------------------------------------------------


def _level_histogram(node_ids: np.ndarray, level_offsets: np.ndarray) -> np.ndarray:
    """How many of ``node_ids`` fall in each level of ``level_offsets``."""
    edges = np.asarray(level_offsets, dtype=np.int64)
    ids = np.sort(np.asarray(node_ids, dtype=np.int64))
    if edges.size < 2:
        if ids.size:
            raise ValueError("selector indices are outside level_offsets")
        return np.zeros(0, dtype=np.int64)
    if ids.size and (int(ids[0]) < int(edges[0]) or int(ids[-1]) >= int(edges[-1])):
        raise ValueError("selector indices are outside level_offsets")
    # ids below each level boundary, differenced: the ids inside each level
    return np.diff(np.searchsorted(ids, edges, side="left")).astype(np.int64)


def _scan_arrays(state: CompiledOperatorState, block_rows) -> dict[str, np.ndarray]:
    """The ``scan_*`` keys: per-block level pairs, shapes, nnz and structural itemsizes,
    and the selector totals (overall and per level)."""
    offsets = np.asarray(state.level_offsets)
    n = len(block_rows)
    table = {name: np.zeros((n,) + shape[1:], dtype=dt) for name, (dt, shape) in _SCAN_SHAPES.items() if shape[0] == "blocks"}
    for i, (d, s, block) in enumerate(block_rows):
        table["scan_block_levels"][i] = (d, s)
        table["scan_block_shapes"][i] = block.shape
        table["scan_block_nnzs"][i] = np.asarray(block.indices).size
        table["scan_block_struct_itemsize"][i] = (np.asarray(block.indices).dtype.itemsize, np.asarray(block.indptr).dtype.itemsize)
    table["scan_selector_nnzs"] = np.asarray((state.sel_mut.nnz, state.sel_miss.nnz), dtype=np.int64)
    table["scan_selector_nnz_by_level"] = np.stack(
        [_level_histogram(sel.indices, offsets) for sel in (state.sel_mut, state.sel_miss)], axis=1).reshape(-1, 2)
    table["scan_metadata_struct_itemsize"] = np.asarray(
        (np.asarray(state.node_perm).dtype.itemsize, np.asarray(state.sample_to_individual).dtype.itemsize), dtype=np.uint8)
    return table


@contextmanager
def _replace_on_success(path: Path):
    """Yield a binary handle on a hidden sibling of ``path``; on a clean exit the
    sibling is flushed to disk and renamed onto ``path``, otherwise it is removed,
    so readers never see a half-written artifact."""
    path.parent.mkdir(parents=True, exist_ok=True)
    sibling = path.with_name(f".{path.name}.{os.getpid()}.{time.monotonic_ns()}.partial")
    fh = open(sibling, "xb")
    try:
        with fh:
            yield fh
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(sibling, path)
    finally:
        if sibling.exists():
            sibling.unlink()


def save_grg_spmv(state: CompiledOperatorState, artifact_path) -> None:
    """Write ``state`` as a format-8 artifact, replacing ``artifact_path`` only once the
    archive is complete on disk (the reference's ``artifact.py:109-204`` contract)."""
    if state.A_blocks is None:
        raise RuntimeError("compiled operator blocks are required for .grg_spmv save")
    if state.init_vector_up_bias is None or state.init_vector_down_bias is None:
        raise RuntimeError("init vector bias cache must be populated before .grg_spmv save")
    started = time.perf_counter()
    archive: dict[str, np.ndarray] = {
        "grg_spmv_magic": np.asarray(GRG_SPMV_FORMAT_MAGIC),
        "grg_spmv_format_version": np.asarray(GRG_SPMV_FORMAT_VERSION, dtype=np.int32),
    }
    for key in _HEADER:
        if key.when is not None and getattr(state, key.when) is None:
            continue
        archive[key.name] = _encode(key.kind, _state_value(state, key.attr))
    for flag, needs in _PRESENCE_FLAGS.items():
        archive[flag] = np.asarray(all(k in archive for k in needs), dtype=bool)
    block_rows = [(d, s, blk) for d, row in enumerate(state.A_blocks) for s, blk in enumerate(row)]
    for d, s, blk in block_rows:
        k_idx, k_ptr = _block_keys(d, s)
        archive[k_idx] = np.asarray(blk.indices)
        archive[k_ptr] = np.asarray(blk.indptr)
    archive.update(_scan_arrays(state, block_rows))
    path = Path(artifact_path)
    with _replace_on_success(path) as fh:
        np.savez(fh, **archive)
    _SCAN_CACHE.clear()
    _LOGGER.debug("save_grg_spmv path=%s blocks=%d elapsed=%.3fs", path, len(block_rows), time.perf_counter() - started)


# ---------------------------------------------------------------------------
# Reading
# ---------------------------------------------------------------------------


def _read_struct(npz, key: str) -> np.ndarray:
    """A structural array (int32 or int64, non-negative). Every structural read goes
    through here, so a test can check which ones a code path touches."""
    arr = np.asarray(npz[key])
    if arr.dtype not in _STRUCT_DTYPES.values():
        raise ValueError(f"unsupported structural dtype for {key}: {arr.dtype}")
    if arr.size and int(arr.min()) < 0:
        raise ValueError(f"structural array {key} must be non-negative")
    return arr


class _Archive:
    """An open artifact whose magic and version were checked on entry."""

    def __init__(self, path) -> None:
        self.path = Path(path)
        self._npz = None

    def __enter__(self) -> "_Archive":
        self._npz = np.load(self.path, allow_pickle=False)
        try:
            self._check_format()
        except BaseException:
            self._npz.close()
            raise
        return self

    def __exit__(self, *exc) -> None:
        self._npz.close()

    def _check_format(self) -> None:
        where = f"unsupported .grg_spmv artifact in {self.path}"
        for key in ("grg_spmv_magic", "grg_spmv_format_version"):
            if key not in self._npz.files:
                raise ValueError(f"{where}: missing {key}")
        magic = str(self.scalar("grg_spmv_magic"))
        if magic != GRG_SPMV_FORMAT_MAGIC:
            raise ValueError(f"{where}: got magic {magic!r}")
        version = int(self.scalar("grg_spmv_format_version"))
        if version != GRG_SPMV_FORMAT_VERSION:
            raise ValueError(f"{where}: got version {version}, expected {GRG_SPMV_FORMAT_VERSION}")

    def has(self, key: str) -> bool:
        return key in self._npz.files

    def scalar(self, key: str):
        return np.asarray(self._npz[key]).item()

    def struct(self, key: str) -> np.ndarray:
        return _read_struct(self._npz, key)

    def array(self, key: str, dtype=None) -> np.ndarray:
        arr = np.asarray(self._npz[key])
        return arr if dtype is None else arr.astype(dtype, copy=False)

    def scan_table(self, key: str, *, blocks: int | None, levels: int) -> np.ndarray:
        """A ``scan_*`` array, checked against its shape in ``_SCAN_SHAPES``."""
        _, shape = _SCAN_SHAPES[key]
        arr = np.asarray(self._npz[key])
        want = tuple({"blocks": blocks, "levels": levels}.get(e, e) for e in shape)
        if arr.ndim != len(want) or any(w is not None and g != w for g, w in zip(arr.shape, want)):
            raise ValueError(f"{key} must have shape {tuple('num_blocks' if e is None else e for e in want)}, got {arr.shape}")
        if arr.size and np.any(arr < 0):
            raise ValueError("scan metadata arrays must be non-negative")
        return arr

    def block_index(self) -> tuple[np.ndarray, np.ndarray]:
        """(level pairs, shapes) of the stored blocks, in stored order."""
        pairs = self.scan_table("scan_block_levels", blocks=None, levels=0)
        shapes = self.scan_table("scan_block_shapes", blocks=pairs.shape[0], levels=0)
        return pairs, shapes

    def decode(self, key: _Key, dtype):
        if key.kind == _COUNT:
            return int(self.scalar(key.name))
        if key.kind == _FLAG:
            return bool(self.scalar(key.name))
        if key.kind == _STRUCT:
            return self.struct(key.name)
        if key.kind == _F64:
            return self.array(key.name, np.float64)
        if key.kind == _I64:
            return self.array(key.name, np.int64)
        if key.kind == _VALUES:
            return self.array(key.name, dtype)
        return self.array(key.name)


_SCAN_CACHE: dict[Path, ArtifactScan] = {}


def _scan(path: Path) -> ArtifactScan:
    with _Archive(path) as ar:
        offsets = ar.struct("level_offsets")
        offsets.setflags(write=False)
        n_levels = max(int(offsets.size) - 1, 0)
        pairs, shapes = ar.block_index()
        nb = int(pairs.shape[0])
        nnzs = ar.scan_table("scan_block_nnzs", blocks=nb, levels=n_levels)
        items = ar.scan_table("scan_block_struct_itemsize", blocks=nb, levels=n_levels)
        sel_total = ar.scan_table("scan_selector_nnzs", blocks=nb, levels=n_levels)
        sel_by_level = ar.scan_table("scan_selector_nnz_by_level", blocks=nb, levels=n_levels).astype(np.int64)
        meta_items = ar.scan_table("scan_metadata_struct_itemsize", blocks=nb, levels=n_levels)
        if np.any(sel_by_level.sum(axis=0) != sel_total):
            raise ValueError("scan selector metadata arrays must have matching nnz totals")
        sel_by_level.setflags(write=False)
        blocks = tuple(
            ArtifactBlockScan(dst_level=int(p[0]), src_level=int(p[1]), shape=(int(sh[0]), int(sh[1])), nnz=int(nz),
                              indices_dtype=_struct_dtype(it[0]), indptr_dtype=_struct_dtype(it[1]))
            for p, sh, nz, it in zip(pairs, shapes, nnzs, items))
        counts = {k.name: int(ar.scalar(k.name)) for k in _HEADER if k.kind == _COUNT}
        return ArtifactScan(
            path=path, **counts,
            has_missing_data=bool(ar.scalar("has_missing_data")),
            has_xtx_bias=bool(ar.scalar("has_xtx_bias")),
            has_individual_coals=bool(ar.scalar("has_individual_coals")),
            level_offsets=offsets,
            level_sizes=tuple(int(v) for v in np.diff(offsets.astype(np.int64))),
            num_levels=n_levels,
            selector_mut_nnz=int(sel_total[0]),
            selector_miss_nnz=int(sel_total[1]),
            selector_nnz_by_level=sel_by_level,
            node_perm_dtype=_struct_dtype(meta_items[0]),
            sample_to_individual_dtype=_struct_dtype(meta_items[1]),
            blocks=blocks,
        )


def scan_grg_spmv(path) -> ArtifactScan:
    """Metadata-only scan: counts, levels, selector totals and every block's shape, nnz and
    structural dtypes, without reading a structural array other than ``level_offsets``.
    Cached per resolved path; ``save_grg_spmv`` clears the cache."""
    resolved = Path(path).resolve()
    hit = _SCAN_CACHE.get(resolved)
    if hit is None:
        hit = _SCAN_CACHE[resolved] = _scan(resolved)
    return hit


def clear_scan_cache() -> None:
    _SCAN_CACHE.clear()


def iter_artifact_blocks(path):
    """Yield every stored level block, one at a time, in stored order."""
    with _Archive(path) as ar:
        pairs, shapes = ar.block_index()
        for (d, s), shape in zip(pairs.tolist(), shapes.tolist()):
            k_idx, k_ptr = _block_keys(d, s)
            yield ArtifactBlock(dst_level=d, src_level=s, shape=(int(shape[0]), int(shape[1])),
                                indices=ar.struct(k_idx), indptr=ar.struct(k_ptr))


def _load_grg_spmv_host(artifact_path, dtype) -> CompiledOperatorState:
    """The state without its level blocks (``A_blocks=None``); init biases in ``dtype``."""
    with _Archive(artifact_path) as ar:
        flags = {flag: bool(ar.scalar(flag)) for flag in _PRESENCE_FLAGS}
        optional = {k for flag, keys in _PRESENCE_FLAGS.items() if flags[flag] for k in keys}
        fields: dict[str, object] = {}
        csr: dict[str, dict[str, np.ndarray]] = {}
        for key in _HEADER:
            if key.when is not None and key.name not in optional:
                fields[key.attr] = None
                continue
            value = ar.decode(key, dtype)
            owner, _, part = key.attr.partition(".")
            if part:
                csr.setdefault(owner, {})[part] = value
            else:
                fields[key.attr] = value
    shape = (fields["num_mutations"], fields["num_nodes"])
    for owner, parts in csr.items():
        fields[owner] = binary_csr_from_parts(indices=parts["indices"], indptr=parts["indptr"], shape=shape)
    return CompiledOperatorState(A_blocks=None, inv_node_perm=_invert_permutation(fields["node_perm"]), **fields)


def load_grg_spmv(artifact_path, dtype) -> CompiledOperatorState:
    """Host state plus every level block as a binary CSR (narrowed dtypes). Level pairs
    that the archive does not store are empty blocks of the right shape."""
    state = _load_grg_spmv_host(artifact_path, dtype)
    sizes = np.diff(np.asarray(state.level_offsets, dtype=np.int64)).tolist()
    grid = [[binary_csr_from_parts(indices=np.empty(0, np.int32), indptr=np.zeros(sizes[d] + 1, np.int32),
                                   shape=(sizes[d], sizes[s])) for s in range(d)] for d in range(len(sizes))]
    oversized = False
    for blk in iter_artifact_blocks(artifact_path):
        csr = binary_csr_from_parts(indices=blk.indices, indptr=blk.indptr, shape=blk.shape, shared_data=True)
        oversized |= (csr.indices.dtype.itemsize < blk.indices.dtype.itemsize
                      or csr.indptr.dtype.itemsize < blk.indptr.dtype.itemsize)
        grid[blk.dst_level][blk.src_level] = csr
    state.A_blocks = grid
    if oversized:
        warnings.warn(f"{artifact_path} stores some structural arrays wider than necessary; "
                      "load_grg_spmv() rebuilt smaller in-memory dtypes, so the artifact uses more disk space than needed",
                      RuntimeWarning, stacklevel=2)
    return state


__all__ = [
    "ArtifactBlock",
    "ArtifactBlockScan",
    "ArtifactScan",
    "GRG_SPMV_FORMAT_MAGIC",
    "GRG_SPMV_FORMAT_VERSION",
    "artifact_path_for_grg",
    "clear_scan_cache",
    "iter_artifact_blocks",
    "load_grg_spmv",
    "save_grg_spmv",
    "scan_grg_spmv",
]
