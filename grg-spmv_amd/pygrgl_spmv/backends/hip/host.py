"""ctypes binding of ``libgrg_host.so`` (``include/grg_host.h``): host preprocessing."""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from ctypes import c_int32, c_int64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parents[2] / "_lib" / "libgrg_host.so"


class GrgHostBlock(ctypes.Structure):
    _fields_ = [("dst_level", c_int32), ("src_level", c_int32), ("nnz", c_int64), ("indptr", c_void_p), ("indices", c_void_p)]


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.grg_host_locality_order.argtypes = [c_int64, c_int64, c_int32, c_void_p, ctypes.POINTER(GrgHostBlock), c_int32,
                                                c_int64, ctypes.c_double, c_void_p]
        lib.grg_host_locality_order.restype = ctypes.c_int
        lib.grg_host_cluster_order.argtypes = [c_int64, c_int32, c_void_p, ctypes.POINTER(GrgHostBlock), c_int32, c_int64,
                                               c_int64, ctypes.c_double, c_int64, c_void_p]
        lib.grg_host_cluster_order.restype = ctypes.c_int
        lib.grg_host_build_tiles.argtypes = [c_int64, c_int32, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                             c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                             c_void_p, c_void_p]
        lib.grg_host_build_tiles.restype = ctypes.c_int
        lib.grg_host_merged_keys.argtypes = [c_int64, c_int32, c_void_p, c_void_p, c_void_p, c_void_p]
        lib.grg_host_merged_keys.restype = ctypes.c_int
        lib.grg_host_build_panels.argtypes = [c_int64, c_int64, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int64, c_int32]
        lib.grg_host_build_panels.restype = c_void_p
        lib.grg_host_panels_sizes.argtypes = [c_void_p, c_void_p]
        lib.grg_host_panels_sizes.restype = ctypes.c_int
        lib.grg_host_panels_copy.argtypes = [c_void_p] * 7
        lib.grg_host_panels_copy.restype = ctypes.c_int
        lib.grg_host_panels_free.argtypes = [c_void_p]
        lib.grg_host_panels_free.restype = None
        lib.grg_host_build_groups.argtypes = [c_int64, c_int64, c_void_p, c_void_p, c_int64, c_int32]
        lib.grg_host_build_groups.restype = c_void_p
        lib.grg_host_groups_sizes.argtypes = [c_void_p, c_void_p]
        lib.grg_host_groups_sizes.restype = ctypes.c_int
        lib.grg_host_groups_copy.argtypes = [c_void_p] * 5
        lib.grg_host_groups_copy.restype = ctypes.c_int
        lib.grg_host_groups_free.argtypes = [c_void_p]
        lib.grg_host_groups_free.restype = None
        lib.grg_host_last_error.argtypes = []
        lib.grg_host_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def _host_blocks(blocks):
    keep = []
    arr = (GrgHostBlock * max(1, len(blocks)))()
    for i, (d, s, ptr, idx) in enumerate(blocks):
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        keep.append((ptr, idx))
        arr[i] = GrgHostBlock(int(d), int(s), int(idx.size), ptr.ctypes.data, idx.ctypes.data)
    return arr, keep


def locality_order(num_nodes: int, num_samples: int, level_offsets: np.ndarray, blocks, *, degree_window: int = 0,
                   degree_min_mean: float = 0.0) -> np.ndarray:
    """``new_pos[old]`` for blocks given as (dst, src, indptr int64, indices int32) tuples.

    ``degree_window > 1`` orders each run of that many consecutive nodes of a level
    (levels above 0 with mean child count >= ``degree_min_mean``) by child count."""
    lib = _lib()
    off = np.ascontiguousarray(level_offsets, dtype=np.int64)
    arr, keep = _host_blocks(blocks)
    out = np.empty(int(num_nodes), dtype=np.int64)
    rc = lib.grg_host_locality_order(int(num_nodes), int(num_samples), int(off.size - 1), off.ctypes.data, arr, len(blocks),
                                     int(degree_window), float(degree_min_mean), out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"grg_host_locality_order failed: status {rc}: {lib.grg_host_last_error().decode()}")
    return out


def cluster_order(new_pos: np.ndarray, level_offsets: np.ndarray, blocks, *, window: int, degree_window: int = 0,
                  min_mean: float = 0.0, min_rows: int = 0) -> np.ndarray:
    """``new_pos`` refined by the min-child order (``grg_host_cluster_order``): inside
    windows of ``window`` consecutive positions of every level above 0 with at least
    ``min_rows`` nodes and a mean child count of at least ``min_mean``, nodes ordered by
    the position of their first child; then, with ``degree_window > 1``, aligned runs of
    that many positions by child count. Returns a new array; the input is not modified."""
    lib = _lib()
    off = np.ascontiguousarray(level_offsets, dtype=np.int64)
    arr, keep = _host_blocks(blocks)
    out = np.array(new_pos, dtype=np.int64, copy=True)
    rc = lib.grg_host_cluster_order(int(out.size), int(off.size - 1), off.ctypes.data, arr, len(blocks), int(window),
                                    int(degree_window), float(min_mean), int(min_rows), out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"grg_host_cluster_order failed: status {rc}: {lib.grg_host_last_error().decode()}")
    return out


@dataclass
class Tiles:
    """Tiles of one sweep configuration (``grg_host_build_tiles``): int32 ``tile_row`` /
    ``tile_src`` (num_tiles + 1), uint32 ``src_list``, uint16 ``lidx`` (per merged
    entry, plus LIDX_PAD zeros so 16-byte reads past the end stay inside), int32
    ``eptr`` (rows + 1), per-level ``level_tile_off`` (L + 1), and per level the most
    staged sources, rows and entries of any staged tile (``level_max_staged``,
    ``level_max_rows``, ``level_max_entries``)."""

    tile_row: np.ndarray
    tile_src: np.ndarray
    src_list: np.ndarray
    lidx: np.ndarray
    eptr: np.ndarray
    level_tile_off: np.ndarray
    level_max_staged: np.ndarray
    level_max_rows: np.ndarray
    level_max_entries: np.ndarray

    @property
    def nbytes(self) -> int:
        return int(sum(a.nbytes for a in (self.tile_row, self.tile_src, self.src_list, self.lidx, self.eptr)))


LIDX_PAD = 8


def build_tiles(level_offsets: np.ndarray, segments, heavy_threshold, *, max_staged: int = 256, max_rows: int = 512,
                max_entries: int = 4096) -> Tiles:
    """Tiles over every level for segments given as (ptr int32[rows + 1], idx int32, source_rows)."""
    lib = _lib()
    off = np.ascontiguousarray(level_offsets, dtype=np.int64)
    L = off.size - 1
    R = int(off[-1])
    nseg = len(segments)
    if not 1 <= nseg <= 3:
        raise ValueError("1 to 3 segments")
    ptrs = [np.ascontiguousarray(p, dtype=np.int32) for p, _, _ in segments]
    idxs = [np.ascontiguousarray(i, dtype=np.int32) for _, i, _ in segments]
    for p in ptrs:
        if p.size != R + 1:
            raise ValueError("segment pointers must cover every row")
    deg = np.zeros(R, dtype=np.int64)
    for p in ptrs:
        deg += np.diff(p.astype(np.int64))
    eptr64 = np.zeros(R + 1, dtype=np.int64)
    np.cumsum(deg, out=eptr64[1:])
    total = int(eptr64[-1])
    if total > np.iinfo(np.int32).max:
        raise NotImplementedError("more than 2^31 - 1 entries in one sweep configuration")
    eptr = eptr64.astype(np.int32)
    src_rows = np.asarray([int(n) for _, _, n in segments], dtype=np.int64)
    heavy = np.ascontiguousarray(heavy_threshold, dtype=np.int64)
    if heavy.size != L:
        raise ValueError("one heavy threshold per level")
    cap_tiles = R + L + 1
    tile_row = np.empty(cap_tiles, dtype=np.int32)
    tile_src = np.empty(cap_tiles, dtype=np.int32)
    src_list = np.empty(max(total, 1), dtype=np.uint32)
    lidx = np.zeros(total + LIDX_PAD, dtype=np.uint16)
    level_tile_off = np.empty(L + 1, dtype=np.int64)
    level_max = np.zeros(max(3 * L, 1), dtype=np.int32)
    pp = (c_void_p * nseg)(*[p.ctypes.data for p in ptrs])
    ip = (c_void_p * nseg)(*[i.ctypes.data for i in idxs])
    rc = lib.grg_host_build_tiles(R, L, off.ctypes.data, nseg, ctypes.cast(pp, c_void_p), ctypes.cast(ip, c_void_p),
                                  src_rows.ctypes.data, eptr.ctypes.data, heavy.ctypes.data, int(max_staged), int(max_rows),
                                  int(max_entries), tile_row.ctypes.data, tile_src.ctypes.data, src_list.ctypes.data, lidx.ctypes.data,
                                  level_tile_off.ctypes.data, level_max.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"grg_host_build_tiles failed: status {rc}: {lib.grg_host_last_error().decode()}")
    nt = int(level_tile_off[-1])
    ns = int(tile_src[nt])
    return Tiles(tile_row=tile_row[: nt + 1].copy(), tile_src=tile_src[: nt + 1].copy(), src_list=src_list[: max(ns, 1)].copy(),
                 lidx=lidx, eptr=eptr, level_tile_off=level_tile_off, level_max_staged=level_max[:L].copy(),
                 level_max_rows=level_max[L:2 * L].copy(), level_max_entries=level_max[2 * L:3 * L].copy())


def merged_keys(eptr: np.ndarray, segments) -> np.ndarray:
    """uint32 (segment << 30) | source row of every entry, rows' entries over all
    segments in segment order at ``eptr`` (``grg_host_merged_keys``)."""
    lib = _lib()
    eptr = np.ascontiguousarray(eptr, dtype=np.int32)
    R = eptr.size - 1
    ptrs = [np.ascontiguousarray(p, dtype=np.int32) for p, _, _ in segments]
    idxs = [np.ascontiguousarray(i, dtype=np.int32) for _, i, _ in segments]
    nseg = len(segments)
    ekey = np.empty(max(int(eptr[-1]), 1), dtype=np.uint32)
    pp = (c_void_p * nseg)(*[p.ctypes.data for p in ptrs])
    ip = (c_void_p * nseg)(*[i.ctypes.data for i in idxs])
    rc = lib.grg_host_merged_keys(R, nseg, ctypes.cast(pp, c_void_p), ctypes.cast(ip, c_void_p), eptr.ctypes.data, ekey.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"grg_host_merged_keys failed: status {rc}: {lib.grg_host_last_error().decode()}")
    return ekey


__all__ = ["LIB_PATH", "LIDX_PAD", "Tiles", "build_tiles", "cluster_order", "locality_order", "merged_keys"]


@dataclass
class Panels:
    """Panels of one level launch (``grg_host_build_panels``): int32 ``pan_pass``
    (panels + 1), ``pass_src`` / ``pass_ent`` (passes + 1), uint16 ``pass_rowoff``
    (passes x (panel_rows + 1)), uint32 ``src_list`` and uint8 ``lidx`` (each padded
    with PANEL_PAD zeros so 16-byte reads past the end stay inside)."""

    panel_rows: int
    max_staged: int
    max_pass_entries: int
    max_passes: int
    pan_pass: np.ndarray
    pass_src: np.ndarray
    pass_ent: np.ndarray
    pass_rowoff: np.ndarray
    src_list: np.ndarray
    lidx: np.ndarray

    @property
    def num_panels(self) -> int:
        return int(self.pan_pass.size - 1)

    @property
    def nbytes(self) -> int:
        return int(sum(a.nbytes for a in (self.pan_pass, self.pass_src, self.pass_ent, self.pass_rowoff, self.src_list,
                                          self.lidx)))


PANEL_PAD = 16


def build_panels(eptr: np.ndarray, keys: np.ndarray, row_begin: int, row_end: int, *, panel_rows: int = 1024,
                 max_staged: int = 255, max_pass_entries: int = 8000, heavy_threshold: int = -1, threads: int = 0) -> Panels:
    """Panels for rows [row_begin, row_end) of entries ``keys[eptr[r]:eptr[r + 1]]``: passes
    of at most ``max_staged`` distinct sources and ``max_pass_entries`` entries."""
    import os

    lib = _lib()
    eptr = np.ascontiguousarray(eptr, dtype=np.int32)
    keys = np.ascontiguousarray(keys).view(np.uint32) if np.asarray(keys).dtype.itemsize == 4 else np.ascontiguousarray(keys, dtype=np.uint32)
    if not 0 <= int(row_begin) <= int(row_end) <= eptr.size - 1:
        raise ValueError("row range outside eptr")
    n = int(threads) if threads > 0 else min(16, os.cpu_count() or 1)
    h = lib.grg_host_build_panels(int(row_begin), int(row_end), eptr.ctypes.data, keys.ctypes.data, int(panel_rows),
                                  int(max_staged), int(max_pass_entries), int(heavy_threshold), n)
    if not h:
        raise RuntimeError(f"grg_host_build_panels failed: {lib.grg_host_last_error().decode()}")
    try:
        sz = np.zeros(6, dtype=np.int64)
        if lib.grg_host_panels_sizes(h, sz.ctypes.data) != 0:
            raise RuntimeError(f"grg_host_panels_sizes failed: {lib.grg_host_last_error().decode()}")
        n_pan, n_pass, n_src, n_ent, max_ent, max_passes = (int(v) for v in sz)
        pan_pass = np.empty(n_pan + 1, dtype=np.int32)
        pass_src = np.empty(n_pass + 1, dtype=np.int32)
        pass_ent = np.empty(n_pass + 1, dtype=np.int32)
        pass_rowoff = np.empty(max(n_pass * (int(panel_rows) + 1), 1), dtype=np.uint16)
        src_list = np.zeros(n_src + PANEL_PAD, dtype=np.uint32)
        lidx = np.zeros(n_ent + PANEL_PAD, dtype=np.uint8)
        if lib.grg_host_panels_copy(h, pan_pass.ctypes.data, pass_src.ctypes.data, pass_ent.ctypes.data, pass_rowoff.ctypes.data,
                                    src_list.ctypes.data, lidx.ctypes.data) != 0:
            raise RuntimeError(f"grg_host_panels_copy failed: {lib.grg_host_last_error().decode()}")
    finally:
        lib.grg_host_panels_free(h)
    return Panels(panel_rows=int(panel_rows), max_staged=int(max_staged), max_pass_entries=max_ent, max_passes=max_passes,
                  pan_pass=pan_pass,
                  pass_src=pass_src, pass_ent=pass_ent, pass_rowoff=pass_rowoff, src_list=src_list, lidx=lidx)


GROUP_ROWS = 16  # rows of one group of the row-group launch (the MFMA tile height)


@dataclass
class Groups:
    """Row groups of one level launch (``grg_host_build_groups``): int64 ``off`` (groups + 1,
    multiples of 4), uint16 ``store`` (groups), uint32 ``key`` and uint16 ``mask`` (entries)."""

    off: np.ndarray
    store: np.ndarray
    key: np.ndarray
    mask: np.ndarray

    @property
    def num_groups(self) -> int:
        return int(self.store.size)

    @property
    def num_entries(self) -> int:
        return int(self.key.size)

    @property
    def nbytes(self) -> int:
        return int(sum(a.nbytes for a in (self.off, self.store, self.key, self.mask)))


def build_groups(eptr: np.ndarray, keys: np.ndarray, row_begin: int, row_end: int, *, heavy_threshold: int = -1,
                 threads: int = 0) -> Groups:
    """Row groups of 16 consecutive rows for rows [row_begin, row_end) of entries
    ``keys[eptr[r]:eptr[r + 1]]``: each group's distinct keys with the mask of rows listing them."""
    import os

    lib = _lib()
    eptr = np.ascontiguousarray(eptr, dtype=np.int32)
    keys = np.ascontiguousarray(keys).view(np.uint32) if np.asarray(keys).dtype.itemsize == 4 else np.ascontiguousarray(keys, dtype=np.uint32)
    if not 0 <= int(row_begin) <= int(row_end) <= eptr.size - 1:
        raise ValueError("row range outside eptr")
    n = int(threads) if threads > 0 else min(16, os.cpu_count() or 1)
    h = lib.grg_host_build_groups(int(row_begin), int(row_end), eptr.ctypes.data, keys.ctypes.data, int(heavy_threshold), n)
    if not h:
        raise RuntimeError(f"grg_host_build_groups failed: {lib.grg_host_last_error().decode()}")
    try:
        sz = np.zeros(2, dtype=np.int64)
        if lib.grg_host_groups_sizes(h, sz.ctypes.data) != 0:
            raise RuntimeError(f"grg_host_groups_sizes failed: {lib.grg_host_last_error().decode()}")
        n_groups, n_ent = int(sz[0]), int(sz[1])
        off = np.empty(n_groups + 1, dtype=np.int64)
        store = np.empty(max(n_groups, 1), dtype=np.uint16)
        key = np.empty(max(n_ent, 1), dtype=np.uint32)
        mask = np.empty(max(n_ent, 1), dtype=np.uint16)
        if lib.grg_host_groups_copy(h, off.ctypes.data, store.ctypes.data, key.ctypes.data, mask.ctypes.data) != 0:
            raise RuntimeError(f"grg_host_groups_copy failed: {lib.grg_host_last_error().decode()}")
    finally:
        lib.grg_host_groups_free(h)
    return Groups(off=off, store=store[:n_groups], key=key[:n_ent], mask=mask[:n_ent])
