"""Binary CSR helpers (reference ``pygrgl_spmv/grg/sparse.py:13-77``).

GRG blocks are 0/1 matrices: only ``indices``/``indptr`` are stored; the value
array is a bool view (optionally one shared read-only ``True`` broadcast).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

_INT32_MAX = int(np.iinfo(np.int32).max)
_ONE_TRUE = np.ones(1, dtype=np.bool_)
_ONE_TRUE.setflags(write=False)


def pick_small_signed_int_dtype(max_value: int) -> np.dtype:
    """int32 when ``max_value`` fits, else int64; negatives are rejected."""
    value = int(max_value)
    if value < 0:
        raise ValueError(f"Structural integer values must be non-negative, got {value}")
    return np.dtype(np.int32) if value <= _INT32_MAX else np.dtype(np.int64)


def finalize_int_array(values, *, label: str, non_negative: bool = True) -> np.ndarray:
    """Narrow an integer array to the smallest of int32/int64 holding its values."""
    arr = np.asarray(values)
    if arr.dtype.kind not in ("i", "u"):
        raise TypeError(f"{label} must be an integer array, got {arr.dtype}")
    if arr.ndim == 0:
        arr = arr.reshape(1)
    if arr.size == 0:
        return arr if arr.dtype == np.dtype(np.int32) else arr.astype(np.int32)
    lo = int(arr.min())
    hi = int(arr.max())
    if non_negative and lo < 0:
        raise ValueError(f"{label} must be non-negative")
    bound = hi if non_negative else max(abs(lo), abs(hi))
    target = pick_small_signed_int_dtype(bound)
    return arr if arr.dtype == target else arr.astype(target)


def binary_csr_from_parts(*, indices, indptr, shape, shared_data: bool = False) -> sp.csr_matrix:
    """A bool-valued CSR from stored ``indices``/``indptr``."""
    shape_tuple = tuple(int(v) for v in shape)
    if len(shape_tuple) != 2:
        raise ValueError(f"CSR shape must have 2 dimensions, got {shape_tuple}")
    idx = np.asarray(indices)
    ptr = np.asarray(indptr)
    if idx.ndim != 1 or ptr.ndim != 1:
        raise ValueError("CSR indices and indptr must be one-dimensional")
    for name, arr in (("indices", idx), ("indptr", ptr)):
        if arr.dtype not in (np.dtype(np.int32), np.dtype(np.int64)):
            raise TypeError(f"{name} must use int32 or int64, got {arr.dtype}")
    nnz = int(idx.size)
    if nnz == 0:
        data = np.empty(0, dtype=np.bool_)
    elif shared_data:
        data = np.broadcast_to(_ONE_TRUE, (nnz,))
    else:
        data = np.ones(nnz, dtype=np.bool_)
    return sp.csr_matrix((data, idx, ptr), shape=shape_tuple)


__all__ = ["binary_csr_from_parts", "finalize_int_array", "pick_small_signed_int_dtype"]
