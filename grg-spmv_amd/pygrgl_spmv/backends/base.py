"""Backend-shared planning helpers.

These keep the byte-accounting and ordering contracts of the reference's
``pygrgl_spmv/backends/base.py:15-205`` (budget items, structural dtype bounds,
``(dst, src)`` level-pair order, per-level selector split, ring-slot reuse links)
so a HIP layout budgets exactly like the Triton/cuSPARSE planners do.
"""

from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Iterator

import numpy as np
import scipy.sparse as sp

from pygrgl_spmv.backends.types import Direction, SparseFormat, StoredMatrix

_INT32_MAX = int(np.iinfo(np.int32).max)


@dataclass(frozen=True)
class BudgetItem:
    """One line of a layout's device-memory budget (reference ``base.py:15-23``)."""

    kind: str
    name: str
    nbytes: int
    artifact_index: int | None = None
    dst_level: int | None = None
    src_level: int | None = None
    slot: int | None = None


def struct_dtype_for_bound(max_value: int) -> np.dtype:
    """Smallest of int32/int64 that holds ``max_value`` (reference ``base.py:33-39``)."""
    value = int(max_value)
    if value < 0:
        raise ValueError(f"structural bound must be non-negative, got {value}")
    return np.dtype(np.int64) if value > _INT32_MAX else np.dtype(np.int32)


def require_struct_dtype(dtype, *, label: str) -> np.dtype:
    dt = np.dtype(dtype)
    if dt not in (np.dtype(np.int32), np.dtype(np.int64)):
        raise TypeError(f"{label} must use int32 or int64, got {dt}")
    return dt


def _fmt_token(fmt) -> str:
    return str(getattr(fmt, "value", fmt)).strip().lower()


def layout_struct_dtypes(fmt, *, nrows: int, ncols: int, nnz: int) -> tuple[np.dtype, np.dtype]:
    """(pointer dtype, index dtype) for a stored block (reference ``base.py:42-51``).

    CSR points into ``nnz`` and indexes ``ncols`` columns; CSC points into ``nnz``
    and indexes ``nrows`` rows; COO holds coordinates bounded by the larger side.
    """
    token = _fmt_token(fmt)
    if token == "csr":
        return struct_dtype_for_bound(max(int(nnz), 0)), struct_dtype_for_bound(max(int(ncols) - 1, 0))
    if token == "csc":
        return struct_dtype_for_bound(max(int(nnz), 0)), struct_dtype_for_bound(max(int(nrows) - 1, 0))
    if token == "coo":
        common = struct_dtype_for_bound(max(int(nrows) - 1, int(ncols) - 1, 0))
        return common, common
    raise ValueError(f"unknown sparse format for structural dtype bounds: {fmt!r}")


def sparse_structure_lengths(fmt, *, nrows: int, ncols: int, nnz: int) -> tuple[int, int]:
    """(pointer length, index length) of a stored block (reference ``base.py:54-62``)."""
    token = _fmt_token(fmt)
    if token == "csr":
        return int(nrows) + 1, int(nnz)
    if token == "csc":
        return int(ncols) + 1, int(nnz)
    if token == "coo":
        return int(nnz), int(nnz)
    raise ValueError(f"unknown sparse format for structural lengths: {fmt!r}")


def sparse_structure_nbytes(fmt, *, nrows: int, ncols: int, nnz: int, struct0_dtype=None, struct1_dtype=None) -> int:
    """Structural bytes of one stored block (reference ``base.py:65-81``)."""
    d0, d1 = layout_struct_dtypes(fmt, nrows=nrows, ncols=ncols, nnz=nnz)
    d0 = d0 if struct0_dtype is None else np.dtype(struct0_dtype)
    d1 = d1 if struct1_dtype is None else np.dtype(struct1_dtype)
    n0, n1 = sparse_structure_lengths(fmt, nrows=nrows, ncols=ncols, nnz=nnz)
    return int(n0 * d0.itemsize + n1 * d1.itemsize)


def copy_struct_checked(dst, values, *, label: str) -> None:
    """Copy a structural array into ``dst`` refusing silent int64->int32 overflow."""
    dst_arr = np.asarray(dst)
    src_arr = np.asarray(values)
    src_dt = require_struct_dtype(src_arr.dtype, label=label)
    dst_dt = require_struct_dtype(dst_arr.dtype, label=label)
    if dst_arr.shape != src_arr.shape:
        raise ValueError(f"{label} shape mismatch: expected {dst_arr.shape}, got {src_arr.shape}")
    if src_dt.itemsize > dst_dt.itemsize and src_arr.size:
        info = np.iinfo(dst_dt)
        if int(src_arr.min()) < int(info.min) or int(src_arr.max()) > int(info.max):
            raise ValueError(f"{label} exceeds {dst_dt} range")
    np.copyto(dst_arr, src_arr, casting="unsafe")


def iter_direction_level_pairs(direction: Direction, height: int) -> Iterator[tuple[int, int, int]]:
    """Yield ``(dst_level, src_level, position)`` in sweep order (reference ``base.py:106-119``).

    UP visits, for each destination level, its sources from the lowest level
    upward. DOWN visits, for each destination level, its sources from the top
    level downward; ``position`` is the source's distance above ``dst`` minus one.
    """
    if height < 0:
        raise ValueError(f"number of levels must be non-negative, got {height}")
    if direction == Direction.UP:
        for dst in range(height):
            for src in range(dst):
                yield dst, src, src
    elif direction == Direction.DOWN:
        for dst in range(height):
            for src in range(height - 1, dst, -1):
                yield dst, src, src - dst - 1
    else:
        raise ValueError(f"unknown direction {direction!r}")


def stored_block_shape(nrows: int, ncols: int, *, store: StoredMatrix) -> tuple[int, int]:
    return (int(nrows), int(ncols)) if store == StoredMatrix.N else (int(ncols), int(nrows))


def materialize_sparse_block(block: sp.spmatrix, *, store: StoredMatrix, fmt: SparseFormat) -> sp.spmatrix:
    """Canonical stored form of ``block`` (or its transpose) in ``fmt`` (reference ``base.py:128-149``)."""
    stored = block if store == StoredMatrix.N else block.T
    if fmt == SparseFormat.CSR:
        out = stored.tocsr()
    elif fmt == SparseFormat.CSC:
        out = stored.tocsc()
    elif fmt == SparseFormat.COO:
        out = stored.tocoo()
        out.sum_duplicates()
        return out
    else:
        raise ValueError(f"unsupported sparse format: {fmt}")
    out.sum_duplicates()
    out.sort_indices()
    return out


def split_selector_by_level(selector: sp.spmatrix, level_offsets) -> list[tuple[np.ndarray, np.ndarray]]:
    """Per level, the selector's (row, level-local column) pairs (reference ``base.py:152-175``)."""
    coo = selector.tocoo()
    cols = np.asarray(coo.col, dtype=np.int64)
    order = np.argsort(cols, kind="stable")
    rows_sorted = np.asarray(coo.row)[order]
    cols_sorted = cols[order]
    row_dt = struct_dtype_for_bound(max(int(selector.shape[0]) - 1, 0))
    offsets = np.asarray(level_offsets, dtype=np.int64)
    cuts = np.searchsorted(cols_sorted, offsets)
    out: list[tuple[np.ndarray, np.ndarray]] = []
    for level in range(offsets.size - 1):
        a, b = int(cuts[level]), int(cuts[level + 1])
        lo, hi = int(offsets[level]), int(offsets[level + 1])
        col_dt = struct_dtype_for_bound(max(hi - lo - 1, 0))
        out.append((rows_sorted[a:b].astype(row_dt, copy=False), (cols_sorted[a:b] - lo).astype(col_dt)))
    return out


def relink_stream_dependencies(ops_by_level, *, direction: Direction):
    """Point each slot-using op at the previous op that used the same ring slot.

    Reference ``base.py:194-205``: the H2D copy of a streamed block must wait
    for the compute of the block that last occupied its slot, in sweep order.
    """
    last: dict[int, tuple[int, int]] = {}
    height = len(ops_by_level)
    order = range(1, height) if direction == Direction.UP else range(height - 2, -1, -1)
    for dst in order:
        for i, op in enumerate(ops_by_level[dst]):
            slot = getattr(op, "slot", None)
            if slot is None:
                continue
            ops_by_level[dst][i] = replace(op, prev_in_slot=last.get(int(slot)))
            last[int(slot)] = (dst, i)
    return ops_by_level


__all__ = [
    "BudgetItem",
    "copy_struct_checked",
    "iter_direction_level_pairs",
    "layout_struct_dtypes",
    "materialize_sparse_block",
    "relink_stream_dependencies",
    "require_struct_dtype",
    "sparse_structure_lengths",
    "sparse_structure_nbytes",
    "split_selector_by_level",
    "stored_block_shape",
    "struct_dtype_for_bound",
]
