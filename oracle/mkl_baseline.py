"""CPU baseline: the reference's MKL sweep, restated -- TEST/BENCH INFRASTRUCTURE ONLY.

Mirrors ``MklRuntime._run`` (reference ``pygrgl_spmv/backends/mkl/backend.py:559-645``):
one persistent MKL handle per non-empty level block (CSR of ``A[d][s]``, values
= ones), ``mkl_sparse_d_mm`` (``mkl_sparse_d_mv`` at k = 1) with alpha = beta = 1,
up with ``SPARSE_OPERATION_NON_TRANSPOSE`` and down with
``SPARSE_OPERATION_TRANSPOSE`` on the same handles (the reference's default
up N/CSR + down T/CSC pair shares storage), mm / mv hints and
``mkl_sparse_optimize`` at setup (``backend.py:496-525``), threads set with
``MKL_Set_Num_Threads_Local`` (``ffi.py:496-498``). Seeding (selector scatter)
and the selector gather are outside the timed sweep, as in the reference's
``scripts/bench/run.py:34-70`` protocol (which times the sweep call only).
Only ``bench.py``'s cpu-baseline leg and tests use this module.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_void_p

import numpy as np

SPARSE_INDEX_BASE_ZERO = 0
SPARSE_OPERATION_NON_TRANSPOSE = 10
SPARSE_OPERATION_TRANSPOSE = 11
SPARSE_MATRIX_TYPE_GENERAL = 20
SPARSE_FILL_MODE_LOWER = 40
SPARSE_DIAG_NON_UNIT = 50
SPARSE_LAYOUT_ROW_MAJOR = 101


class _Descr(ctypes.Structure):
    _fields_ = [("type", c_int), ("mode", c_int), ("diag", c_int)]


def _load_mkl():
    for cand in (os.environ.get("MKL_RT"), "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so", "libmkl_rt.so.2", "libmkl_rt.so.1"):
        if not cand:
            continue
        try:
            # DEEPBIND: libtorch exports MKL symbols of its own; MKL must bind to itself.
            lib = ctypes.CDLL(cand, mode=os.RTLD_LOCAL | getattr(os, "RTLD_DEEPBIND", 0))
        except OSError:
            continue
        lib.MKL_Set_Interface_Layer.argtypes = [c_int]
        lib.MKL_Set_Interface_Layer.restype = c_int
        lib.MKL_Set_Interface_Layer(0)  # LP64: 32-bit MKL_INT
        # GNU OpenMP threading: torch already loaded libgomp, and a second
        # OpenMP runtime (Intel's) in the same process crashes.
        lib.MKL_Set_Threading_Layer.argtypes = [c_int]
        lib.MKL_Set_Threading_Layer.restype = c_int
        lib.MKL_Set_Threading_Layer(3)  # MKL_THREADING_GNU
        lib.mkl_sparse_d_create_csr.argtypes = [POINTER(c_void_p), c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]
        lib.mkl_sparse_d_mm.argtypes = [c_int, c_double, c_void_p, _Descr, c_int, c_void_p, c_int, c_int, c_double, c_void_p, c_int]
        lib.mkl_sparse_d_mv.argtypes = [c_int, c_double, c_void_p, _Descr, c_void_p, c_double, c_void_p]
        lib.mkl_sparse_set_mm_hint.argtypes = [c_void_p, c_int, _Descr, c_int, c_int, c_int]
        lib.mkl_sparse_set_mv_hint.argtypes = [c_void_p, c_int, _Descr, c_int]
        lib.mkl_sparse_optimize.argtypes = [c_void_p]
        lib.mkl_sparse_destroy.argtypes = [c_void_p]
        lib.MKL_Set_Num_Threads_Local.argtypes = [c_int]
        lib.MKL_Set_Num_Threads_Local.restype = c_int
        return lib, cand
    return None, None


class MklSweep:
    """Per-block MKL sweep over an oracle-loaded GRG (``oracle.grg_oracle.OracleGRG``)."""

    def __init__(self, g, *, max_k: int, n_threads: int = 0, expected_calls: int = 1000):
        lib, path = _load_mkl()
        if lib is None:
            raise RuntimeError("libmkl_rt not found")
        self.lib, self.path = lib, path
        self.g = g
        self.n_threads = int(n_threads) if n_threads else (os.cpu_count() or 1)
        self.descr = _Descr(SPARSE_MATRIX_TYPE_GENERAL, SPARSE_FILL_MODE_LOWER, SPARSE_DIAG_NON_UNIT)
        self._keep = []
        self.handles = {}
        for (d, s), b in g.blocks.items():
            ptr = np.ascontiguousarray(b.indptr, dtype=np.int32)
            idx = np.ascontiguousarray(b.indices, dtype=np.int32)
            vals = np.ones(idx.size, dtype=np.float64)
            h = c_void_p()
            st = lib.mkl_sparse_d_create_csr(byref(h), SPARSE_INDEX_BASE_ZERO, b.shape[0], b.shape[1],
                                             ptr.ctypes.data, ptr[1:].ctypes.data, idx.ctypes.data, vals.ctypes.data)
            if st != 0:
                raise RuntimeError(f"mkl_sparse_d_create_csr failed: status {st}")
            for op in (SPARSE_OPERATION_NON_TRANSPOSE, SPARSE_OPERATION_TRANSPOSE):
                lib.mkl_sparse_set_mv_hint(h, op, self.descr, expected_calls)
                if max_k > 1:
                    lib.mkl_sparse_set_mm_hint(h, op, self.descr, SPARSE_LAYOUT_ROW_MAJOR, max_k, expected_calls)
            lib.mkl_sparse_optimize(h)
            self._keep.append((ptr, idx, vals))
            self.handles[(d, s)] = h

    def close(self):
        for h in self.handles.values():
            self.lib.mkl_sparse_destroy(h)
        self.handles = {}

    def _apply(self, h, op, src: np.ndarray, dst: np.ndarray, k: int):
        if k == 1:
            st = self.lib.mkl_sparse_d_mv(op, 1.0, h, self.descr, src.ctypes.data, 1.0, dst.ctypes.data)
        else:
            st = self.lib.mkl_sparse_d_mm(op, 1.0, h, self.descr, SPARSE_LAYOUT_ROW_MAJOR, src.ctypes.data, k,
                                          src.strides[0] // 8, 1.0, dst.ctypes.data, dst.strides[0] // 8)
        if st != 0:
            raise RuntimeError(f"mkl_sparse_d_{'mv' if k == 1 else 'mm'} failed: status {st}")

    def sweep(self, values: np.ndarray, direction: str) -> None:
        """In-place level sweep on node values (N, k) float64 C-contiguous."""
        off = self.g.level_offsets
        L = self.g.height
        k = values.shape[1]
        prev = self.lib.MKL_Set_Num_Threads_Local(self.n_threads)
        try:
            if direction == "up":
                for d in range(1, L):
                    dst = values[off[d] : off[d + 1]]
                    for s in range(d):
                        h = self.handles.get((d, s))
                        if h is not None:
                            self._apply(h, SPARSE_OPERATION_NON_TRANSPOSE, values[off[s] : off[s + 1]], dst, k)
            else:
                for d in range(L - 2, -1, -1):
                    dst = values[off[d] : off[d + 1]]
                    for s in range(L - 1, d, -1):
                        h = self.handles.get((s, d))
                        if h is not None:
                            self._apply(h, SPARSE_OPERATION_TRANSPOSE, values[off[s] : off[s + 1]], dst, k)
        finally:
            self.lib.MKL_Set_Num_Threads_Local(prev)


def mkl_matmul(g, sweep: MklSweep, x: np.ndarray, direction: str, *, miss: np.ndarray | None = None) -> np.ndarray:
    """``BoundGRG.matmul`` without init modes or by-individual columns, with the level
    sweep done by ``sweep`` (the reference's MklRuntime arithmetic,
    ``mkl/backend.py:559-645``) and the selectors by SciPy: up seeds the sample rows
    with ``x`` (k, S) and returns ``sel_mut @ V`` (k, M); down seeds with
    ``sel_mut^T y`` (+ ``sel_miss^T miss``) and returns the sample rows (k, S)."""
    xi = np.ascontiguousarray(np.asarray(x, dtype=np.float64).T)
    k = xi.shape[1]
    if direction == "up":
        v = np.zeros((g.num_nodes, k))
        v[: g.num_samples] = xi
        sweep.sweep(v, "up")
        return np.asarray(g.sel_mut @ v).T.copy()
    seed = g.sel_mut.T @ xi
    if miss is not None:
        seed = seed + g.sel_miss.T @ np.asarray(miss, dtype=np.float64).T
    v = np.ascontiguousarray(seed, dtype=np.float64)
    sweep.sweep(v, "down")
    return v[: g.num_samples].T.copy()


__all__ = ["MklSweep", "mkl_matmul"]
