"""ctypes binding of ``libgrg_hip.so`` (declared in ``include/grg_hip.h``).

Error convention as the reference's FFIs (``mkl/ffi.py:237-240``,
``cusparse/ffi.py:75-78``): a non-zero status raises
``RuntimeError("HIP <fn> failed: status <rc>: <message>")``. There is no
fallback: if the library is missing the import of this module fails loudly.
"""

from __future__ import annotations

import ctypes
from ctypes import c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parents[2] / "_lib" / "libgrg_hip.so"
ABI_VERSION = 17
BARRIER_WORDS = 40960  # GRG_HIP_BARRIER_WORDS: uint32 words of the persistent barrier buffer
STATIC_SCRATCH_DOUBLES = 1 << 22
F64 = 0
F32C = 1
FLAG_XCD_REMAP = 2
FLAG_PARTIALS = 4
FLAG_TILED = 8
FLAG_MERGED = 16
FLAG_STREAM = 32
FLAG_PANEL = 64
FLAG_XCD_RUNS = 128
XCD_RUN_SHIFT = 8
FLAG_GROUP_KEYS = 4096
FLAG_GROUPS = 8192
GROUP_ROWS = 16
TILE_CHUNK_BYTES = 128
TILE_MAX_STAGED = 1024
MAX_SEGS = 3
MAX_MAPS = 2


class GrgHipSeg(ctypes.Structure):
    _fields_ = [("ptr", c_void_p), ("idx", c_void_p), ("src", c_void_p), ("ld", c_int64), ("ptr_base", c_int64)]


class GrgHipMap(ctypes.Structure):
    _fields_ = [("map", c_void_p), ("src", c_void_p), ("ld", c_int64), ("row_lo", c_int64), ("row_hi", c_int64)]


class GrgHipGatherArgs(ctypes.Structure):
    _fields_ = [
        ("dtype", c_int32),
        ("accumulate", c_int32),
        ("row_begin", c_int64),
        ("row_end", c_int64),
        ("k", c_int32),
        ("cpl", c_int32),
        ("kl", c_int32),
        ("es", c_int32),
        ("unroll", c_int32),
        ("rows_per_group", c_int32),
        ("flags", c_int32),
        ("out", c_void_p),
        ("ld_out", c_int64),
        ("out_base", c_int64),
        ("nseg", c_int32),
        ("nmap", c_int32),
        ("seg", GrgHipSeg * MAX_SEGS),
        ("map", GrgHipMap * MAX_MAPS),
        ("row_scale", c_void_p),
        ("scale_base", c_int64),
        ("col_vec", c_void_p),
        ("heavy_threshold", c_int64),
        ("heavy_rows", c_void_p),
        ("num_heavy", c_int64),
        ("split_ptr", c_void_p),
        ("split_row", c_void_p),
        ("split_items", c_int64),
        ("split_chunk", c_int64),
        ("partials", c_void_p),
        ("out_scale", c_void_p),
        ("out_scale_base", c_int64),
        ("tile_row", c_void_p),
        ("tile_src", c_void_p),
        ("src_list", c_void_p),
        ("eptr", c_void_p),
        ("lidx", c_void_p),
        ("num_tiles", c_int64),
        ("tile_lds_rows", c_int64),
        ("tile_max_rows", c_int64),
        ("tile_max_entries", c_int64),
        ("ekey", c_void_p),
        ("stream_blocks", c_int64),
        ("pan_pass", c_void_p),
        ("pass_src", c_void_p),
        ("pass_ent", c_void_p),
        ("pass_rowoff", c_void_p),
        ("pan_src", c_void_p),
        ("pan_lidx", c_void_p),
        ("num_panels", c_int64),
        ("panel_rows", c_int32),
        ("panel_staged", c_int32),
        ("panel_max_entries", c_int32),
        ("panel_max_passes", c_int32),
        ("grp_off", c_void_p),
        ("grp_store", c_void_p),
        ("grp_key", c_void_p),
        ("grp_mask", c_void_p),
        ("num_groups", c_int64),
    ]


class GrgHipJob(ctypes.Structure):
    _fields_ = [("args", GrgHipGatherArgs), ("block_begin", c_int64), ("gx", c_int64), ("gy", c_int64), ("n_light", c_int64),
                ("g_shift", c_int32), ("kl_shift", c_int32)]


class HipLibrary:
    def __init__(self, path: Path = LIB_PATH):
        if not Path(path).exists():
            raise RuntimeError(
                f"HIP extension {path} is not built; run __graft_entry__.build() -- the HIP backend has no CPU fallback"
            )
        # torch ships its own HIP runtime under the same soname (libamdhip64.so.7): load it
        # first, so that this library binds to it and torch's streams, events and device
        # memory are the ones the kernels see (loaded the other way round, the system
        # runtime would shadow torch's and torch would find no GPU)
        import torch  # noqa: F401

        lib = ctypes.CDLL(str(path))
        lib.grg_hip_gather.argtypes = [ctypes.POINTER(GrgHipGatherArgs), c_void_p]
        lib.grg_hip_gather.restype = ctypes.c_int
        lib.grg_hip_gather_batch.argtypes = [ctypes.POINTER(GrgHipGatherArgs), c_int32, c_void_p]
        lib.grg_hip_gather_batch.restype = ctypes.c_int
        lib.grg_hip_gather_batch_static.argtypes = [ctypes.POINTER(GrgHipGatherArgs), c_int32, c_void_p]
        lib.grg_hip_gather_batch_static.restype = ctypes.c_int
        lib.grg_hip_last_error.argtypes = []
        lib.grg_hip_last_error.restype = ctypes.c_char_p
        lib.grg_hip_abi_version.argtypes = []
        lib.grg_hip_abi_version.restype = c_int32
        lib.grg_hip_device_count.argtypes = [ctypes.POINTER(c_int32)]
        lib.grg_hip_device_count.restype = ctypes.c_int
        lib.grg_hip_stream_device.argtypes = [c_void_p, ctypes.POINTER(c_int32)]
        lib.grg_hip_stream_device.restype = ctypes.c_int
        lib.grg_hip_stream_capturing.argtypes = [c_void_p, ctypes.POINTER(c_int32)]
        lib.grg_hip_stream_capturing.restype = ctypes.c_int
        lib.grg_hip_jobs_build.argtypes = [ctypes.POINTER(GrgHipGatherArgs), c_void_p, c_int32, c_int32,
                                           ctypes.POINTER(GrgHipJob), c_void_p, c_void_p]
        lib.grg_hip_jobs_build.restype = ctypes.c_int
        lib.grg_hip_jobs_max_grid.argtypes = [c_int32, c_void_p, ctypes.POINTER(c_int32)]
        lib.grg_hip_jobs_max_grid.restype = ctypes.c_int
        lib.grg_hip_jobs_run.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                                         c_void_p, c_void_p, c_void_p]
        lib.grg_hip_jobs_run.restype = ctypes.c_int
        lib.grg_hip_row_weighted_sum.argtypes = [c_int32, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                                                 c_int32, c_void_p]
        lib.grg_hip_row_weighted_sum.restype = ctypes.c_int
        version = int(lib.grg_hip_abi_version())
        if version != ABI_VERSION:
            raise RuntimeError(f"libgrg_hip.so ABI version {version} != expected {ABI_VERSION}; rebuild the extension")
        self.lib = lib
        self.path = Path(path)

    def _check(self, rc: int, fn: str) -> None:
        if int(rc) != 0:
            msg = self.lib.grg_hip_last_error().decode("utf-8", errors="replace")
            raise RuntimeError(f"HIP {fn} failed: status {int(rc)}: {msg}")

    def gather(self, args: GrgHipGatherArgs, stream_ptr: int) -> None:
        self._check(self.lib.grg_hip_gather(ctypes.byref(args), c_void_p(int(stream_ptr))), "grg_hip_gather")

    def gather_batch(self, array, n: int, stream_ptr: int) -> None:
        self._check(self.lib.grg_hip_gather_batch(array, c_int32(int(n)), c_void_p(int(stream_ptr))), "grg_hip_gather_batch")

    def gather_batch_static(self, array, n: int, stream_ptr: int) -> None:
        """``grg_hip_gather_batch`` with the split-row partials in the library's static per-GPU
        array instead of each descriptor's ``partials`` (no allocation on this path)."""
        self._check(self.lib.grg_hip_gather_batch_static(array, c_int32(int(n)), c_void_p(int(stream_ptr))),
                    "grg_hip_gather_batch_static")

    def jobs_build(self, args_list, phase_of):
        """Job table of launch descriptors (grg_hip_jobs_build): (jobs ctypes array,
        phase_job_off int32[nphase + 1], phase_blocks int64[nphase])."""
        import numpy as np

        n = len(args_list)
        phase = np.ascontiguousarray(phase_of, dtype=np.int32)
        nphase = int(phase.max()) + 1 if n else 0
        arr = (GrgHipGatherArgs * max(n, 1))(*args_list)
        jobs = (GrgHipJob * max(n, 1))()
        job_off = np.zeros(nphase + 1, dtype=np.int32)
        blocks = np.zeros(max(nphase, 1), dtype=np.int64)
        self._check(self.lib.grg_hip_jobs_build(arr, phase.ctypes.data, n, nphase, jobs, job_off.ctypes.data, blocks.ctypes.data),
                    "grg_hip_jobs_build")
        return jobs, job_off, blocks[:nphase]

    def jobs_max_grid(self, dtype: int, stream_ptr: int) -> int:
        out = c_int32(0)
        self._check(self.lib.grg_hip_jobs_max_grid(int(dtype), c_void_p(int(stream_ptr)), ctypes.byref(out)), "grg_hip_jobs_max_grid")
        return int(out.value)

    def jobs_run(self, jobs_dev: int, job_off_dev: int, blocks_dev: int, blocks_host, dtype: int, nphase: int,
                 persistent: bool, grid: int, barrier_dev: int, status_dev: int, stream_ptr: int) -> None:
        self._check(self.lib.grg_hip_jobs_run(c_void_p(int(jobs_dev)), c_void_p(int(job_off_dev)), c_void_p(int(blocks_dev)),
                                              c_void_p(blocks_host.ctypes.data), int(dtype), int(nphase), int(bool(persistent)),
                                              int(grid), c_void_p(int(barrier_dev)), c_void_p(int(status_dev)),
                                              c_void_p(int(stream_ptr))), "grg_hip_jobs_run")

    def row_weighted_sum(self, dtype: int, x_ptr: int, rows: int, cols: int, ld: int, w_ptr: int, out_ptr: int,
                         partials_ptr: int, n_blocks: int, stream_ptr: int) -> None:
        self._check(self.lib.grg_hip_row_weighted_sum(int(dtype), c_void_p(int(x_ptr)), int(rows), int(cols), int(ld),
                                                      c_void_p(int(w_ptr)), c_void_p(int(out_ptr)), c_void_p(int(partials_ptr)),
                                                      int(n_blocks), c_void_p(int(stream_ptr))), "grg_hip_row_weighted_sum")

    def device_count(self) -> int:
        out = c_int32(0)
        self._check(self.lib.grg_hip_device_count(ctypes.byref(out)), "grg_hip_device_count")
        return int(out.value)

    def stream_device(self, stream_ptr: int) -> int:
        out = c_int32(-1)
        self._check(self.lib.grg_hip_stream_device(c_void_p(int(stream_ptr)), ctypes.byref(out)), "grg_hip_stream_device")
        return int(out.value)

    def stream_capturing(self, stream_ptr: int) -> bool:
        out = c_int32(0)
        self._check(self.lib.grg_hip_stream_capturing(c_void_p(int(stream_ptr)), ctypes.byref(out)), "grg_hip_stream_capturing")
        return bool(out.value)


_LIB: HipLibrary | None = None


def hip_library() -> HipLibrary:
    global _LIB
    if _LIB is None:
        _LIB = HipLibrary()
    return _LIB


__all__ = ["ABI_VERSION", "F32C", "F64", "GrgHipGatherArgs", "GrgHipJob", "GrgHipMap", "GrgHipSeg", "HipLibrary", "LIB_PATH", "hip_library"]
