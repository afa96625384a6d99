"""ctypes binding of ``libgrg_host.so`` (``include/grg_host.h``): host preprocessing."""

from __future__ import annotations

import ctypes
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
        lib.grg_host_locality_order.argtypes = [c_int64, c_int64, c_int32, c_void_p, ctypes.POINTER(GrgHostBlock), c_int32, c_void_p]
        lib.grg_host_locality_order.restype = ctypes.c_int
        lib.grg_host_last_error.argtypes = []
        lib.grg_host_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def locality_order(num_nodes: int, num_samples: int, level_offsets: np.ndarray, blocks) -> np.ndarray:
    """``new_pos[old]`` for blocks given as (dst, src, indptr int64, indices int32) tuples."""
    lib = _lib()
    off = np.ascontiguousarray(level_offsets, dtype=np.int64)
    keep = []
    arr = (GrgHostBlock * max(1, len(blocks)))()
    for i, (d, s, ptr, idx) in enumerate(blocks):
        ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        keep.append((ptr, idx))
        arr[i] = GrgHostBlock(int(d), int(s), int(idx.size), ptr.ctypes.data, idx.ctypes.data)
    out = np.empty(int(num_nodes), dtype=np.int64)
    rc = lib.grg_host_locality_order(int(num_nodes), int(num_samples), int(off.size - 1), off.ctypes.data, arr, len(blocks), out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"grg_host_locality_order failed: status {rc}: {lib.grg_host_last_error().decode()}")
    return out


__all__ = ["LIB_PATH", "locality_order"]
