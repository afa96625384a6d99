"""``HipRuntime``: runtime-owned HIP execution of the GRG matmul.

Counterpart of the reference's ``TritonRuntime`` (``triton/backend.py:609-1225``)
with the same protocol towards ``BoundGRG``: everything is allocated in
``__enter__`` and released in ``__exit__`` (also when entering fails);
``.grgs`` raises before entry; ``_call_scope`` rejects re-entrant calls;
``_prepare_matmul_cuda`` returns a prepared op whose ``input`` / ``output`` /
``miss_input`` / ``miss_output`` / ``init_vector`` / ``init_matrix`` tensors
are transposed views of one shared arena (so one direction's output is the
other direction's input); ``op()`` performs no allocation, host sync or
device read-back and, when every block is resident, can be captured into a
hipGraph.

Execution differs from the reference by design:

* the resident blocks of each destination level are merged on the device, at
  ``__enter__``, into one CSR over global node ids -- parent-major (children
  of each node) for the up sweep, child-major (parents of each node) for the
  down sweep -- so each level is ONE launch that writes each row once, with
  no atomics;
* the seed pass (input, by-individual expansion, selector scatter, missing
  data, init vector / matrix / XTX) is fused into the level launches as
  extra gather terms, and the write-out (selector gather, missing output,
  endpoint bias, by-individual reduce, emit_all_nodes permutation) is one
  more launch;
* a whole resident sweep is a single ``grg_hip_gather_batch`` call on the
  runtime's root stream; streamed (non-resident) blocks are copied from
  pinned host memory into ring slots on per-slot copy streams, each copy
  waiting only for the previous compute that used its slot, and are then
  accumulated into their destination level.
"""

from __future__ import annotations

from contextlib import contextmanager
import ctypes
import hashlib
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from pygrgl_spmv.backends.base import copy_struct_checked, relink_stream_dependencies
from pygrgl_spmv.backends.hip import ffi, tunecache
from pygrgl_spmv.backends.hip.layout import (PANEL_PASS_ENTRIES, PANEL_ROWS, PANEL_STAGED, SPLIT_CHUNK, SPLIT_MIN, HipLayout,
                                             enabled_max_k, side_io_rows)
from pygrgl_spmv.backends.types import Direction, InitMode
from pygrgl_spmv.grg import BoundGRG, _CudaMatmulSpec
from pygrgl_spmv.grg.artifact import _load_grg_spmv_host, iter_artifact_blocks


def _torch():
    import torch

    return torch


# --------------------------------------------------------------------------
# Launch-shape selection
# --------------------------------------------------------------------------


def _pow2_ceil(v: int) -> int:
    return 1 if v <= 1 else 1 << (int(v) - 1).bit_length()


def _pow2_floor(v: float) -> int:
    return 1 if v < 2 else 1 << (int(v).bit_length() - 1)


KERNEL_UNROLL = 8  # gathers in flight per lane and round in the kernel
TILE_MAX_STAGED = 128   # staged source rows of one tile (LDS rows)
TILE_MAX_ROWS = 256     # rows of one tile
TILE_MAX_ENTRIES = 2048  # gather entries of one tile
PANEL_MIN_MEAN = 3.0    # levels whose light rows average fewer entries get no panels
PANEL_MIN_ROWS = 4096   # ... nor levels with fewer rows
PANEL_BLOCK = 1024            # threads of a panel workgroup
PANEL_MAX_PASS_BYTES = 8 * PANEL_BLOCK   # entry positions of one pass (with 32 bytes of slack) the panel kernel stages
PANEL_MAX_PASSES = 2048       # passes of one panel (their entry offsets sit in LDS)
GROUP_MIN_MEAN = 3.0    # levels whose light rows average fewer entries get no row groups
GROUP_MIN_ROWS = 256    # ... nor levels with fewer rows
GROUP_MAX_FILL = 0.9    # ... nor levels whose groups keep more than this share of the entries
DEGREE_WINDOW = 4096    # nodes per degree-sorted run of the device node order
DEGREE_MIN_MEAN = 6.0
CLUSTER_WINDOW = 0          # nodes per min-child-ordered window of the device node order (0: off)
CLUSTER_DEGREE_WINDOW = 0   # aligned runs inside those windows ordered by degree (0: none)
CLUSTER_MIN_ROWS = 65536    # levels with fewer nodes keep the locality order
# Run lengths (log2 of light blocks) the tuner tries for GRG_HIP_FLAG_XCD_RUNS.
XCD_RUN_LOGS = (4, 5, 6, 8, 9, 10, 12)
# The tuner's last stage: the fastest few distinct launches re-timed, interleaved, in
# this many rounds, for launches of at least this many milliseconds.
TUNE_FINALISTS = 4
TUNE_RETIME_ROUNDS = 3
TUNE_RETIME_MIN_MS = 0.05
XCD_RUN_BITS = 15 << ffi.XCD_RUN_SHIFT


def xcd_runs_flags(log_run: int) -> int:
    """Launch flags for XCD runs of 2^log_run light blocks (GRG_HIP_FLAG_XCD_RUNS)."""
    if not 0 <= int(log_run) <= 15:
        raise ValueError("log_run must be in [0, 15]")
    return ffi.FLAG_XCD_RUNS | (int(log_run) << ffi.XCD_RUN_SHIFT)   # levels with a lower mean up-sweep degree keep the plain locality order


def column_shapes(dtype_code: int, k: int, lds) -> list[tuple[int, int]]:
    """Admissible (cpl, kl): cpl-wide vector columns (up to 32 bytes per lane) that
    divide k and every leading dimension, kl lanes covering k (<= 64 columns per chunk)."""
    elem = 8 if dtype_code == ffi.F64 else 4
    out = []
    for cpl in (8, 4, 2, 1):
        if cpl * elem > 32 or k % cpl or any(int(ld) % cpl for ld in lds):
            continue
        out.append((cpl, min(_pow2_ceil(-(-k // cpl)), 64 // cpl)))
    return out or [(1, min(_pow2_ceil(k), 64))]


def launch_shape(dtype_code: int, k: int, lds, mean_degree: float) -> tuple[int, int, int]:
    """Default (cpl, kl, es) before tuning: 16-byte columns, and edge lanes so one
    round of KERNEL_UNROLL gathers per lane covers a typical row."""
    shapes = column_shapes(dtype_code, k, lds)
    elem = 8 if dtype_code == ffi.F64 else 4
    cpl, kl = next(((c, l) for c, l in shapes if c * elem <= 16), shapes[-1])
    es = max(1, min(_pow2_floor(max(mean_degree, 1.0) / KERNEL_UNROLL), 64 // kl))
    return cpl, kl, es


TILE_LDS_LIMIT = 160 * 1024  # LDS per CU on gfx950


def tile_lds_bytes(dtype_code: int, cpl: int, kl: int, staged_rows: int, max_rows: int, max_entries: int) -> int:
    """LDS of one tiled workgroup: staged rows of kl * cpl columns plus a padding of at
    least 16 bytes and one lane vector, the tile's row pointers and entry indices (each
    part rounded to 16 bytes), and the 2 KiB reduction buffer."""
    elem = 8 if dtype_code == ffi.F64 else 4
    pad = cpl if cpl * elem >= 16 else 16 // elem

    def r16(b):
        return (int(b) + 15) // 16 * 16

    return r16(int(staged_rows) * (kl * cpl + pad) * elem) + r16((int(max_rows) + 1) * 4) + r16((int(max_entries) + 16) * 2) + 8 * 4 * 64


def panel_lds_bytes(dtype_code: int, cpl: int, kl: int, staged_rows: int, panel_rows: int, max_entries: int,
                    max_passes: int) -> int:
    """LDS of one panel workgroup: the panel's int32 pass source and entry offsets, then
    two pass buffers, each 256 staged rows (as in a tile; the last one zero, the row of
    the padding entries), the uint16 end offsets of the panel's rows and
    PANEL_MAX_PASS_BYTES of entry positions (each part rounded to 16 bytes).
    (``staged_rows`` is at most 255 and ``max_entries`` bounded by PANEL_MAX_PASS_BYTES.)"""
    elem = 8 if dtype_code == ffi.F64 else 4
    pad = cpl if cpl * elem >= 16 else 16 // elem

    def r16(b):
        return (int(b) + 15) // 16 * 16

    buffer = r16(256 * (kl * cpl + pad) * elem) + r16(int(panel_rows) * 2) + PANEL_MAX_PASS_BYTES
    return r16(2 * (int(max_passes) + 1) * 4) + 2 * buffer


def panel_rows_per_group(panel_rows: int, kl: int) -> int:
    """Rows per lane group of a panel launch with lane groups of ``kl`` lanes (1024 / kl
    groups per workgroup), or 0 if ``panel_rows`` cannot be split that way (1, 2 or 4
    rows per group, no more than kl)."""
    groups = PANEL_BLOCK // int(kl)
    if int(panel_rows) % groups:
        return 0
    rpg = int(panel_rows) // groups
    return rpg if rpg in (1, 2, 4) and rpg <= int(kl) else 0


@dataclass(frozen=True)
class _LevelSchedule:
    """Heavy rows of one level launch: rows of gather degree > ``heavy_threshold``
    are produced by a workgroup each and listed at ``heavy_offset`` in the
    schedule buffer; ``mean_degree`` is the mean over the other rows. Heavy
    rows of degree > split_min are split: ``num_items`` chunks of
    ``split_chunk`` entries, with item offsets per heavy row at
    ``split_ptr_offset`` and the heavy index of each item at
    ``split_row_offset`` (all int32 entries of the schedule buffer)."""

    heavy_offset: int
    num_heavy: int
    heavy_threshold: int
    mean_degree: float
    split_ptr_offset: int = 0
    split_row_offset: int = 0
    num_items: int = 0
    split_chunk: int = 0


HEAVY_MEAN_FACTOR = 8.0  # rows above this multiple of their level's mean degree are heavy
TUNE_INFLIGHT_BYTES = 256  # tuner: gathered bytes in flight per lane (256: 64 registers)


def level_schedule(deg: np.ndarray, lo: int, hi: int, *, heavy_min: int, cursor: int,
                   split_min: int = 0, split_chunk: int = 0):
    """Schedule for rows [lo, hi) of one sweep configuration (``deg`` = gather degree per row).

    Rows above max(heavy_min, HEAVY_MEAN_FACTOR x the level's mean degree) get a 256-thread
    workgroup of their own, so one long row does not hold a wave of short ones;
    rows above ``split_min`` (when > 0) are further cut into chunks of
    ``split_chunk`` entries. Returns the schedule and its int32 table
    (heavy rows, then split item offsets and item rows when any row is split),
    which the caller stores at ``cursor``.
    """
    seg = deg[lo:hi]
    mean = float(seg.mean()) if seg.size else 0.0
    heavy = int(max(heavy_min, HEAVY_MEAN_FACTOR * mean))
    if split_min > 0:
        heavy = min(heavy, int(split_min))  # rows to split are always heavy rows
    heavy_idx = np.flatnonzero(seg > heavy)
    heavy_rows = (heavy_idx + lo).astype(np.int32)
    light = seg[seg <= heavy]
    light_mean = max(1.0, float(light.mean())) if light.size else 1.0
    n_heavy = int(heavy_rows.size)
    if split_min > 0 and n_heavy:
        hdeg = seg[heavy_idx].astype(np.int64)
        items = np.where(hdeg > split_min, -(-hdeg // split_chunk), 0)
        if items.any():
            split_ptr = np.zeros(n_heavy + 1, dtype=np.int64)
            np.cumsum(items, out=split_ptr[1:])
            n_items = int(split_ptr[-1])
            split_row = np.repeat(np.arange(n_heavy, dtype=np.int64), items)
            table = np.concatenate([heavy_rows, split_ptr.astype(np.int32), split_row.astype(np.int32)])
            return _LevelSchedule(cursor, n_heavy, heavy, light_mean, cursor + n_heavy, cursor + 2 * n_heavy + 1,
                                  n_items, int(split_chunk)), table
    return _LevelSchedule(cursor, n_heavy, heavy, light_mean), heavy_rows


# --------------------------------------------------------------------------
# Device-side artifact state
# --------------------------------------------------------------------------


@dataclass
class _StreamStep:
    dst_level: int
    src_level: int
    slot: int
    host_indices: object  # pinned torch int32
    host_indptr: object   # pinned torch int32
    row_lo: int
    row_hi: int
    src_offset: int
    nnz: int
    prev_in_slot: tuple[int, int] | None = None


@dataclass
class _HipArtifact:
    path: Path
    state: object
    level_offsets: np.ndarray
    dev: dict = field(default_factory=dict)           # name -> device tensor
    sched: dict = field(default_factory=dict)         # (config, level) -> _LevelSchedule
    up_stream: list = field(default_factory=list)     # per level: list[_StreamStep]
    down_stream: list = field(default_factory=list)
    up_events: list = field(default_factory=list)     # per level: list[(copy_done, compute_done)]
    down_events: list = field(default_factory=list)
    max_items: int = 0                                # largest split-item count of one level launch
    tiles: dict = field(default_factory=dict)         # config -> per-level tile offsets and LDS maxima
    panels: dict = field(default_factory=dict)        # (config, level) -> (panels, most entries of a pass, most passes of a panel)
    groups: dict = field(default_factory=dict)        # (config, level) -> (row groups, group entries)


@dataclass
class _Buffers:
    """Value buffers a prepared sweep writes: node state (rows of ``ld_state``
    values), the dense arenas ``io0`` / ``io1`` / ``aux`` (2-D, k columns at most)
    and the FP64 partial rows of split rows. The runtime's own arena serves the
    ordinary ops; a batched sweep gives every GRG a private set."""

    state: object
    ld_state: int
    io0: object
    io1: object
    aux: object
    partials: object


def _i32(values, label: str) -> np.ndarray:
    arr = np.asarray(values)
    out = np.empty(arr.shape, dtype=np.int32)
    copy_struct_checked(out, arr.astype(np.int64, copy=False) if arr.dtype not in (np.int32, np.int64) else arr, label=label)
    return out


def _csr_transpose(indptr, indices, n_rows: int, n_cols: int):
    m = sp.csr_matrix((np.ones(len(indices), dtype=np.int8), np.asarray(indices, dtype=np.int64), np.asarray(indptr, dtype=np.int64)), shape=(n_rows, n_cols))
    t = m.T.tocsr()
    t.sort_indices()
    return np.asarray(t.indptr, dtype=np.int64), np.asarray(t.indices, dtype=np.int64)


def _remapped_block(indptr, indices, new_pos: np.ndarray, off: np.ndarray, d: int, s: int, *, transpose: bool):
    """A block's CSR in the device node order: rows of level d and columns of level s
    (or, with ``transpose``, rows of level s and columns of level d), level-local, sorted."""
    rows_d = int(off[d + 1] - off[d])
    rows_s = int(off[s + 1] - off[s])
    ptr = np.asarray(indptr, dtype=np.int64)
    r = np.repeat(np.arange(rows_d, dtype=np.int64), np.diff(ptr))
    c = np.asarray(indices, dtype=np.int64)
    r_new = new_pos[r + off[d]] - off[d]
    c_new = new_pos[c + off[s]] - off[s]
    if transpose:
        m = sp.csr_matrix((np.ones(r.size, dtype=np.int8), (c_new, r_new)), shape=(rows_s, rows_d))
    else:
        m = sp.csr_matrix((np.ones(r.size, dtype=np.int8), (r_new, c_new)), shape=(rows_d, rows_s))
    m.sort_indices()
    return np.asarray(m.indptr, dtype=np.int64), np.asarray(m.indices, dtype=np.int64)


class _HipPreparedMatmul:
    """Prepared op (reference ``_TritonPreparedMatmul``, ``triton/backend.py:564-606``)."""

    def __init__(self, runtime: "HipRuntime", artifact: _HipArtifact, spec: _CudaMatmulSpec) -> None:
        self._runtime = runtime
        self._artifact = artifact
        self._spec = spec
        k = int(spec.k)
        st = artifact.state
        self._input_internal = runtime._io0[: spec.input_cols, :k]
        self.input = self._input_internal.T
        if spec.emit_all_nodes:
            self._output_internal = runtime._aux[: st.num_nodes, :k]
        else:
            self._output_internal = runtime._io0[: spec.output_cols, :k]
        self.output = self._output_internal.T
        if spec.use_miss:
            miss_view = runtime._io1[: st.num_mutations, :k]
            if spec.direction == Direction.DOWN:
                self._miss_input_internal = miss_view
                self.miss_input = miss_view.T
            else:
                self._miss_output_internal = miss_view
                self.miss_output = miss_view.T
        if spec.init_mode == InitMode.VECTOR:
            self._init_vector_internal = runtime._io1[0, :k]
            self.init_vector = self._init_vector_internal
        if spec.init_mode == InitMode.MATRIX:
            self._init_matrix_internal = runtime._io1[: st.num_nodes, :k]
            self.init_matrix = self._init_matrix_internal.T
        self._steps = runtime._build_steps(artifact, spec)
        runtime._flush_tune_cache()

    def __enter__(self) -> "_HipPreparedMatmul":
        return self

    def __exit__(self, exc_type, exc, tb) -> None:
        return None

    @property
    def capturable(self) -> bool:
        """True when ``op()`` can be captured into a hipGraph: every block of the
        sweep is resident. Sweeps that stream blocks through the pinned ring run
        eagerly only (capturing them is refused with a RuntimeError)."""
        return not any(it[0] == "stream" for it in self._steps)

    def set_output_scale(self, scale) -> None:
        """Multiply every output row by ``scale[row]`` inside the write-out launch
        (no extra pass). For the up direction without emit_all_nodes or missing-data
        output; ``scale`` is a device tensor of ``output_cols`` values in the
        runtime's dtype, kept alive by the op. ``None`` removes the scale. Used by
        the standardized-genotype operator (no reference counterpart)."""
        spec = self._spec
        if spec.direction != Direction.UP or spec.emit_all_nodes or spec.use_miss:
            raise ValueError("set_output_scale() applies to up ops without emit_all_nodes or missing-data output")
        kind, arr, n = self._steps[-1]
        if kind != "batch":
            raise RuntimeError("internal error: the write-out is not the last batched launch")
        if scale is None:
            arr[n - 1].out_scale, arr[n - 1].out_scale_base = 0, 0
            self._out_scale = None
            return
        torch = _torch()
        if not isinstance(scale, torch.Tensor) or scale.device != self._runtime.device:
            raise TypeError("scale must be a device tensor on the runtime's device")
        if scale.dtype != self._runtime._torch_dtype() or scale.dim() != 1 or scale.numel() != int(spec.output_cols):
            raise ValueError(f"scale must be a 1-D {self._runtime._torch_dtype()} tensor of {int(spec.output_cols)} values")
        scale = scale.contiguous()
        arr[n - 1].out_scale, arr[n - 1].out_scale_base = scale.data_ptr(), 0
        self._out_scale = scale

    def __call__(self) -> None:
        with self._runtime._call_scope():
            self._call_prelocked()

    def _call_prelocked(self) -> None:
        self._runtime._execute(self._artifact, self._spec, self._steps)


class HipBatchedSweep:
    """Several sweeps (typically one per GRG of ``runtime.grgs``, each with its own k)
    run together from one job table: phase p holds the p-th launch of every sweep,
    so a level of all GRGs is one kernel (``persistent=False``: one launch per
    phase) or the whole batch is ONE persistent launch whose workgroups meet at a
    grid-wide barrier between phases (``persistent=True``).

    No reference counterpart: the reference runs ``runtime.grgs`` one GRG after the
    other. Every sweep writes private buffers, so ``entries[i]`` has its own
    ``inputs[i]`` / ``outputs[i]`` (and ``miss_inputs[i]``, ``init_vectors[i]``...)
    views, laid out as the prepared op's (k x cols, transposed views of row-major
    arenas). Tiled level launches run in their per-row form inside job tables.
    Streamed blocks (layouts under a VRAM budget) are supported by the grouped form:
    the phases run in segments, and between two segments the blocks streamed at that
    point of every sweep are copied into their ring slots (pinned host -> device on
    the slot's copy stream, after the last launch that read the slot) and
    accumulated, as in a single op; such a batch runs eagerly only. The persistent
    form needs every block resident. Calling the object runs the batch on the
    runtime's caller stream (capturable into a hipGraph when every block is
    resident); ``check()`` synchronizes and raises if a persistent barrier wait ran
    past its bound."""

    def __init__(self, runtime: "HipRuntime", entries, *, persistent: bool, grid: int | None) -> None:
        torch = _torch()
        self._runtime = runtime
        self.persistent = bool(persistent)
        self.specs = []
        self.inputs, self.outputs, self.miss_inputs, self.miss_outputs = [], [], [], []
        self.init_vectors, self.init_matrices = [], []
        self._buffers = []
        launches, phase_of = [], []
        streamed = []  # (phase count before it, step, args): blocks streamed between phases
        dev = runtime.device
        tdt = runtime._torch_dtype()
        items_max = max([a.max_items for a in runtime._artifacts] + [1])
        with torch.cuda.device(dev):
            for e in entries:
                e = dict(e)
                grg = e.pop("grg")
                spec = grg._prepare_cuda_spec(direction=e.pop("direction"), k=int(e.pop("k")),
                                              emit_all_nodes=bool(e.pop("emit_all_nodes", False)),
                                              by_individual=bool(e.pop("by_individual", False)),
                                              init_mode=e.pop("init_mode", "none"), use_miss=bool(e.pop("use_miss", False)))
                if e:
                    raise TypeError(f"unknown batched sweep entry keys: {sorted(e)}")
                if spec.direction == Direction.UP and runtime.layout.pair.plan_up is None:
                    raise ValueError("UP plan is not configured")
                if spec.direction == Direction.DOWN and runtime.layout.pair.plan_down is None:
                    raise ValueError("DOWN plan is not configured")
                art = runtime._artifacts[int(grg._artifact_index)]
                st = art.state
                k = int(spec.k)
                io0_rows = max(int(spec.input_cols), int(spec.output_cols), 1)
                io1_rows = 0
                if spec.use_miss:
                    io1_rows = st.num_mutations
                if spec.init_mode == InitMode.VECTOR:
                    io1_rows = max(io1_rows, 1)
                if spec.init_mode == InitMode.MATRIX:
                    io1_rows = max(io1_rows, st.num_nodes)
                bufs = _Buffers(state=torch.zeros((st.num_nodes * k,), dtype=tdt, device=dev), ld_state=k,
                                io0=torch.zeros((io0_rows, k), dtype=tdt, device=dev),
                                io1=torch.zeros((io1_rows, k), dtype=tdt, device=dev) if io1_rows else None,
                                aux=torch.zeros((st.num_nodes, k), dtype=tdt, device=dev) if spec.emit_all_nodes else None,
                                partials=torch.zeros((items_max * k,), dtype=torch.float64, device=dev))
                steps = runtime._build_steps(art, spec, bufs)
                runtime._flush_tune_cache()
                if self.persistent and any(it[0] != "batch" for it in steps):
                    raise ValueError("a persistent batched sweep needs every block of its GRG resident; "
                                     "run a layout with streamed blocks with persistent=False")
                p = 0
                for it in steps:
                    if it[0] == "stream":
                        streamed.append((p, it[1], it[2]))
                        continue
                    _, arr, n = it
                    for i in range(n):
                        a = ffi.GrgHipGatherArgs()
                        ctypes.pointer(a)[0] = arr[i]
                        if a.flags & ffi.FLAG_GROUPS:
                            # job tables run the per-row light path: the row-group shape's 16 column
                            # lanes become one row's lane group, one edge lane, 8 gathers in flight
                            a.es, a.unroll, a.rows_per_group = 1, 8, 1
                        a.flags &= ~(ffi.FLAG_TILED | ffi.FLAG_STREAM | ffi.FLAG_PANEL | ffi.FLAG_XCD_RUNS | XCD_RUN_BITS |
                                     ffi.FLAG_GROUP_KEYS | ffi.FLAG_GROUPS)
                        launches.append(a)
                        phase_of.append(p)
                        p += 1
                self.specs.append(spec)
                self._buffers.append(bufs)
                self.inputs.append(bufs.io0[: spec.input_cols, :k].T)
                out_internal = bufs.aux[: st.num_nodes, :k] if spec.emit_all_nodes else bufs.io0[: spec.output_cols, :k]
                self.outputs.append(out_internal.T)
                miss = bufs.io1[: st.num_mutations, :k].T if spec.use_miss else None
                self.miss_inputs.append(miss if spec.use_miss and spec.direction == Direction.DOWN else None)
                self.miss_outputs.append(miss if spec.use_miss and spec.direction == Direction.UP else None)
                self.init_vectors.append(bufs.io1[0, :k] if spec.init_mode == InitMode.VECTOR else None)
                self.init_matrices.append(bufs.io1[: st.num_nodes, :k].T if spec.init_mode == InitMode.MATRIX else None)
            order = sorted(range(len(launches)), key=lambda i: phase_of[i])
            jobs, job_off, blocks = runtime._lib.jobs_build([launches[i] for i in order], [phase_of[i] for i in order])
            self.num_jobs = len(launches)
            self.num_phases = int(blocks.size)
            raw = np.frombuffer(bytes(jobs), dtype=np.uint8)
            self._jobs_dev = torch.from_numpy(raw.copy()).to(dev)
            self._job_off_dev = torch.from_numpy(job_off).to(dev)
            self._blocks_dev = torch.from_numpy(blocks.copy() if blocks.size else np.zeros(1, np.int64)).to(dev)
            self._blocks_host = np.ascontiguousarray(blocks if blocks.size else np.zeros(1, np.int64))
            # streamed blocks, in phase order (a stable sort keeps each sweep's own order),
            # each with its own copy / compute events
            streamed.sort(key=lambda t: t[0])
            self._streamed = [(p, step, args, torch.cuda.Event(), torch.cuda.Event()) for p, step, args in streamed]
            self._stream_launch = torch.cuda.Event() if streamed else None
            self._barrier = torch.zeros((ffi.BARRIER_WORDS,), dtype=torch.int32, device=dev)
            self._status = torch.zeros((1,), dtype=torch.int32, device=dev)
            max_grid = runtime._lib.jobs_max_grid(runtime._dtype_code, int(runtime._root_stream.cuda_stream))
            want = int(self._blocks_host.max()) if self.num_phases else 1
            self.grid = max(1, min(max_grid, want if grid is None else int(grid)))
            if grid is not None and not 1 <= int(grid) <= max_grid:
                raise ValueError(f"grid must be in [1, {max_grid}] (co-resident workgroups of the persistent kernel)")
        self.max_grid = max_grid

    @property
    def capturable(self) -> bool:
        """Whether the batch may be captured into a hipGraph: False when it streams blocks."""
        return not self._streamed

    def _run_phases(self, p0: int, p1: int) -> None:
        """Phases [p0, p1) of the grouped form (one launch per phase), or all of them
        (persistent form), on the root stream."""
        rt = self._runtime
        if p1 <= p0:
            return
        job_off = self._job_off_dev.data_ptr() + 4 * p0
        blocks = self._blocks_dev.data_ptr() + 8 * p0
        rt._lib.jobs_run(self._jobs_dev.data_ptr(), job_off, blocks, self._blocks_host[p0:p1], rt._dtype_code, p1 - p0,
                         self.persistent, self.grid, self._barrier.data_ptr(), self._status.data_ptr(),
                         int(rt._root_stream.cuda_stream))

    def __call__(self) -> None:
        torch = _torch()
        rt = self._runtime
        with rt._call_scope():
            with rt._caller_root_scope():
                if not self._streamed:
                    self._run_phases(0, self.num_phases)
                    return
                if rt._lib.stream_capturing(int(rt._caller_stream.cuda_stream)):
                    raise RuntimeError("HIP batched sweep: a batch that streams blocks through the pinned ring cannot be "
                                       "captured into a hipGraph")
                root = rt._root_stream
                root_ptr = int(root.cuda_stream)
                # every slot's first copy waits for the work enqueued before this batch
                with torch.cuda.stream(root):
                    self._stream_launch.record(root)
                last_use = {}
                done = 0
                for p, step, args, copy_done, compute_done in self._streamed:
                    self._run_phases(done, p)
                    done = max(done, p)
                    prev = last_use.get(step.slot, self._stream_launch)
                    rt._copy_block_to_slot(step, prev, copy_done)
                    root.wait_event(copy_done)
                    rt._launch_streamed_block(args, root_ptr)
                    with torch.cuda.stream(root):
                        compute_done.record(root)
                    last_use[step.slot] = compute_done
                self._run_phases(done, self.num_phases)

    def tune_grid(self, candidates=None, *, reps: int = 3) -> dict:
        """Time the persistent launch at each candidate grid (default: a quarter, half,
        three quarters and all of the co-resident workgroups) and keep the fastest. A
        larger grid keeps more gathers in flight per phase; a smaller one makes every
        barrier cheaper. The batch runs ``reps + 1`` times per candidate and its
        outputs overwrite the input views (they share storage): set the inputs again
        afterwards. Returns {grid: best milliseconds}."""
        torch = _torch()
        if not self.persistent:
            raise ValueError("tune_grid applies to a persistent batched sweep")
        if candidates is None:
            candidates = sorted({max(1, self.max_grid * q // 4) for q in (1, 2, 3, 4)})
        rt = self._runtime
        times = {}
        for grid in candidates:
            if not 1 <= int(grid) <= self.max_grid:
                raise ValueError(f"grid must be in [1, {self.max_grid}]")
            self.grid = int(grid)
            best = float("inf")
            for _ in range(int(reps) + 1):
                with torch.cuda.device(rt.device):
                    s = rt._caller_stream   # the batch joins the runtime's caller stream on both sides
                    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    t0.record(s)
                    self()
                    t1.record(s)
                t1.synchronize()
                best = min(best, t0.elapsed_time(t1))
            self.check()
            times[int(grid)] = best
        self.grid = min(times, key=times.get)
        return times

    def barrier_costs(self, *, barriers: int = 64, reps: int = 5) -> dict:
        """Fixed costs of the persistent kernel at this sweep's grid, in microseconds:
        ``one_phase_us``, a persistent run of one empty phase (counter reset, launch
        and registration; no workgroup touches a job), and ``barrier_us``, one grid
        barrier (runs of ``barriers + 1`` empty phases against runs of one). Each run
        is timed with HIP events on the launch stream; best of ``reps``."""
        torch = _torch()
        rt = self._runtime
        dev = rt.device

        def best(nphase: int) -> float:
            off = torch.zeros((nphase + 1,), dtype=torch.int32, device=dev)
            blocks = torch.zeros((nphase,), dtype=torch.int64, device=dev)
            host = np.zeros(nphase, dtype=np.int64)
            times = []
            with rt._caller_root_scope():
                s = rt._root_stream
                for _ in range(reps + 1):
                    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    t0.record(s)
                    rt._lib.jobs_run(self._jobs_dev.data_ptr(), off.data_ptr(), blocks.data_ptr(), host, rt._dtype_code,
                                     nphase, True, self.grid, self._barrier.data_ptr(), self._status.data_ptr(),
                                     int(s.cuda_stream))
                    t1.record(s)
                    t1.synchronize()
                    times.append(t0.elapsed_time(t1))
            self.check()
            return min(times[1:]) * 1e3

        one = best(1)
        return {"one_phase_us": one, "barrier_us": (best(barriers + 1) - one) / barriers}

    def check(self) -> None:
        """Synchronize; raise if any persistent barrier wait ran past its bound."""
        if int(self._status.item()) != 0:
            self._status.zero_()
            raise RuntimeError("HIP batched sweep: a grid barrier wait ran past its bound; the results are invalid")


class HipRuntime:
    """Runtime-owned HIP execution with resident or streamed sparse blocks."""

    def __init__(self, layout: HipLayout, *, autotune: bool = True, split_min: int = SPLIT_MIN,
                 split_chunk: int = SPLIT_CHUNK, fixed_shape: dict | None = None, tiles: bool = True,
                 tile_max_staged: int = TILE_MAX_STAGED, tile_max_rows: int = TILE_MAX_ROWS,
                 tile_max_entries: int = TILE_MAX_ENTRIES, degree_window: int = DEGREE_WINDOW,
                 degree_min_mean: float = DEGREE_MIN_MEAN, cluster_window: int = CLUSTER_WINDOW,
                 cluster_degree_window: int = CLUSTER_DEGREE_WINDOW, cluster_min_rows: int = CLUSTER_MIN_ROWS,
                 panels: bool = False, panel_min_mean: float = PANEL_MIN_MEAN,
                 panel_min_rows: int = PANEL_MIN_ROWS, panel_rows: int = PANEL_ROWS, panel_staged: int = PANEL_STAGED,
                 groups: bool = True,
                 group_min_mean: float = GROUP_MIN_MEAN,
                 group_min_rows: int = GROUP_MIN_ROWS, group_max_fill: float = GROUP_MAX_FILL,
                 tune_cache="default") -> None:
        """``autotune``: time the admissible launch shapes of every level launch when an
        op is prepared (otherwise a fixed default shape). ``split_min`` /
        ``split_chunk``: rows of gather degree above ``split_min`` are summed in
        chunks of ``split_chunk`` entries by separate workgroups (the layout's
        scratch bound assumes the defaults). ``fixed_shape``: launch fields
        (``unroll``, ``rows_per_group``, ``es``, ``flags``, ``cpl``, ``kl``) forced on every launch
        instead of tuning, for tests of each kernel variant (a ``flags`` value with
        ``ffi.FLAG_TILED`` applies to the level launches that have tiles). ``tiles``:
        build the tile tables of the LDS-staged level launches (when the layout
        reserved them) and let the tuner choose them per level; ``tile_max_staged``
        / ``tile_max_rows`` / ``tile_max_entries`` bound the staged sources, rows and
        gather entries of one tile. ``degree_window`` / ``degree_min_mean``: the
        device node order sorts each run of ``degree_window`` consecutive nodes of a
        level (levels whose mean gather degree in the up sweep is at least
        ``degree_min_mean``) by degree, so the rows a wave sums in lockstep have
        similar lengths (0 keeps the plain locality order). ``cluster_window`` (> 1 turns it
        on): then, on levels with at least ``cluster_min_rows`` nodes and that mean degree,
        windows of ``cluster_window`` consecutive nodes are ordered by the position of each
        node's first child, so rows that share it sit side by side (a workgroup's rows gather
        more of the same sources), and aligned runs of ``cluster_degree_window`` nodes inside
        them by degree (``grg_host_cluster_order``). ``panels`` (off by default; only
        under a layout without a VRAM budget): with the tile tables, also build the panels
        of the panel launches (``panel_rows`` rows each: 256,
        512 or 1024; levels whose light rows average at least ``panel_min_mean`` entries,
        with at least ``panel_min_rows`` rows; passes of at most ``panel_staged`` sources)
        and let the tuner choose them. ``groups`` (FP64 values, with the tile tables):
        build the row groups of the row-group launches (levels with at least ``group_min_rows``
        rows whose light rows average at least ``group_min_mean`` entries, kept when a level's
        group entries are at most ``group_max_fill`` of its entries) and let the tuner choose
        them. ``tune_cache``: directory of the on-disk cache of
        tuned launch shapes (``tunecache``; "default": ``$GRG_HIP_TUNE_CACHE`` or
        ``~/.cache/pygrgl_spmv/hip_tune``; None: no disk cache)."""
        if int(degree_window) < 0:
            raise ValueError("degree_window must be >= 0")
        if int(split_min) < 1 or int(split_chunk) < 1:
            raise ValueError("split_min and split_chunk must be >= 1")
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("HIP runtime requires a ROCm GPU (torch.cuda.is_available() is False)")
        self._lib = ffi.hip_library()
        self.layout = layout
        self.device = torch.device("cuda", int(layout.device))
        self.stream_ptr = int(layout.stream_ptr)
        self.autotune = bool(autotune) and not fixed_shape
        allowed = {"unroll", "rows_per_group", "es", "flags", "cpl", "kl", "stream_blocks"}
        if fixed_shape and set(fixed_shape) - allowed:
            raise ValueError(f"fixed_shape keys must be among {sorted(allowed)}")
        self.fixed_shape = dict(fixed_shape or {})
        self.split_min = int(split_min)
        self.split_chunk = int(split_chunk)
        if not 1 <= int(tile_max_staged) <= ffi.TILE_MAX_STAGED or int(tile_max_rows) < 1 or int(tile_max_entries) < 1:
            raise ValueError(f"tile_max_staged must be in [1, {ffi.TILE_MAX_STAGED}], tile_max_rows and tile_max_entries >= 1")
        self.use_tiles = bool(tiles)
        self.tile_max_staged = int(tile_max_staged)
        self.tile_max_rows = int(tile_max_rows)
        self.tile_max_entries = int(tile_max_entries)
        self.use_panels = bool(panels)
        self.use_groups = bool(groups)
        self.group_min_mean = float(group_min_mean)
        self.group_min_rows = int(group_min_rows)
        self.group_max_fill = float(group_max_fill)
        self.panel_min_mean = float(panel_min_mean)
        self.panel_min_rows = int(panel_min_rows)
        if int(panel_rows) not in (256, 512, 1024):
            raise ValueError("panel_rows must be 256, 512 or 1024")
        self.panel_rows = int(panel_rows)
        if not 1 <= int(panel_staged) <= PANEL_STAGED:
            raise ValueError(f"panel_staged must be in [1, {PANEL_STAGED}]")
        if panels and layout.budgeted:
            raise ValueError("panels=True builds panel tables beyond the layout's reservations: it needs a layout "
                             "without a VRAM budget (vram_budget_bytes=0)")
        self.panel_staged = int(panel_staged)
        self.degree_window = int(degree_window)
        self.degree_min_mean = float(degree_min_mean)
        if int(cluster_window) < 0 or int(cluster_degree_window) < 0 or int(cluster_min_rows) < 0:
            raise ValueError("cluster_window, cluster_degree_window and cluster_min_rows must be >= 0")
        self.cluster_window = int(cluster_window)
        self.cluster_degree_window = int(cluster_degree_window)
        self.cluster_min_rows = int(cluster_min_rows)
        self._dtype_code = ffi.F64 if np.dtype(layout.dtype) == np.float64 else ffi.F32C
        self._tune_dir = tunecache.default_directory() if tune_cache == "default" else (
            None if tune_cache is None else Path(tune_cache))
        self._reset_owned_state()

    # ---- lifecycle -------------------------------------------------------

    def _reset_owned_state(self) -> None:
        self.stream = None
        self._caller_stream = None
        self._root_stream = None
        self._caller_to_root_event = None
        self._root_to_caller_event = None
        self._launch_event = None
        self._slot_copy_streams = []
        self._slot_indices = []
        self._slot_indptr = []
        self._state = None
        self._io0 = None
        self._io1 = None
        self._aux = None
        self._partials = None
        self._artifacts = ()
        self._grgs = ()
        self._tune_cache = {}
        self._tune_ctx = {}
        self.tune_log = {}
        self._sweep_timer = None
        self._entered = False
        self._active_call = False

    @property
    def grgs(self) -> tuple:
        if not self._entered:
            raise RuntimeError("HipRuntime must be entered before accessing grgs")
        return self._grgs

    def _torch_dtype(self):
        torch = _torch()
        return torch.float64 if np.dtype(self.layout.dtype) == np.float64 else torch.float32

    def _torch_caller_stream(self):
        return self._caller_stream

    def __enter__(self) -> "HipRuntime":
        torch = _torch()
        try:
            lay = self.layout
            max_k = enabled_max_k(lay.pair, lay.requirements)
            io0_rows = max(lay.max_num_samples, lay.max_num_mutations)
            io1_rows = side_io_rows(lay.pair, lay.requirements, max_num_mutations=lay.max_num_mutations, max_num_nodes=lay.max_num_nodes)
            aux_rows = max(lay.max_num_nodes, lay.max_num_samples, lay.max_num_mutations, lay.max_selector_nnz)
            tdt = self._torch_dtype()
            with torch.cuda.device(self.device):
                self._caller_stream = torch.cuda.get_stream_from_external(self.stream_ptr, device=self.device)
                self.stream = self._caller_stream
                self._root_stream = torch.cuda.Stream(device=self.device)
                self._caller_to_root_event = torch.cuda.Event()
                self._root_to_caller_event = torch.cuda.Event()
                self._launch_event = torch.cuda.Event()
                self._slot_copy_streams = [torch.cuda.Stream(device=self.device) for _ in range(lay.allocated_ring_buffer_size)]
                for slot in lay.slot_plans:
                    if slot.indices_dtype != np.dtype(np.int32) or slot.indptr_dtype != np.dtype(np.int32):
                        raise NotImplementedError("HIP ring slots hold int32 structure; a block exceeds the int32 range")
                self._slot_indices = [torch.zeros((max(1, s.indices_len),), dtype=torch.int32, device=self.device) for s in lay.slot_plans]
                self._slot_indptr = [torch.zeros((max(1, s.indptr_len),), dtype=torch.int32, device=self.device) for s in lay.slot_plans]
                self._state = torch.zeros((lay.max_num_nodes * max_k,), dtype=tdt, device=self.device)
                self._io0 = torch.zeros((io0_rows, max_k), dtype=tdt, device=self.device)
                self._io1 = None if io1_rows == 0 else torch.zeros((io1_rows, max_k), dtype=tdt, device=self.device)
                self._aux = torch.zeros((aux_rows, max_k), dtype=tdt, device=self.device)
            states = tuple(_load_grg_spmv_host(a.path, lay.dtype) for a in lay.artifacts)
            arts = []
            for ai, (al, st) in enumerate(zip(lay.artifacts, states)):
                arts.append(self._build_artifact(ai, al, st))
            self._artifacts = tuple(arts)
            if self.autotune and self._tune_dir is not None:
                for al, art in zip(lay.artifacts, arts):
                    self._tune_ctx[art.path] = tunecache.TuneContext(self._tune_dir, self._tune_context_fields(al, art))
            items = max([a.max_items for a in arts] + [1])
            with torch.cuda.device(self.device):
                # FP64 partial rows of split rows, reused by every level launch in stream order
                self._partials = torch.zeros((items * max_k,), dtype=torch.float64, device=self.device)
            torch.cuda.synchronize(self.device)
            self._grgs = tuple(BoundGRG(self, i, a.state, a.path, lay.device) for i, a in enumerate(self._artifacts))
            self._entered = True
            return self
        except BaseException:
            self._reset_owned_state()
            raise

    def __exit__(self, exc_type, exc, tb) -> None:
        if self._entered:
            _torch().cuda.synchronize(self.device)
        self._reset_owned_state()

    @contextmanager
    def _call_scope(self):
        if not self._entered:
            raise RuntimeError("HipRuntime must be entered before matmul")
        if self._active_call:
            raise RuntimeError("concurrent runtime.grgs calls are not supported")
        self._active_call = True
        try:
            yield
        finally:
            self._active_call = False

    @contextmanager
    def _caller_root_scope(self):
        torch = _torch()
        with torch.cuda.device(self.device):
            with torch.cuda.stream(self._caller_stream):
                self._caller_to_root_event.record(self._caller_stream)
            self._root_stream.wait_event(self._caller_to_root_event)
            try:
                yield
            finally:
                with torch.cuda.stream(self._root_stream):
                    self._root_to_caller_event.record(self._root_stream)
                self._caller_stream.wait_event(self._root_to_caller_event)

    @contextmanager
    def sweep_timer(self):
        """Record a pair of timing events on the launch stream around every op run
        inside the block; yields the list of (start, end) events (read them after a
        synchronize). Used to time the kernels on the stream they run on."""
        if self._sweep_timer is not None:
            raise RuntimeError("sweep_timer() is already active")
        self._sweep_timer = []
        try:
            yield self._sweep_timer
        finally:
            self._sweep_timer = None

    def prepare_batched_sweep(self, entries, *, persistent: bool = True, grid: int | None = None) -> HipBatchedSweep:
        """Sweeps of several GRGs run from one job table (see ``HipBatchedSweep``).
        ``entries``: dicts with ``grg`` (one of ``self.grgs``), ``direction``, ``k``
        and optionally ``emit_all_nodes``, ``by_individual``, ``init_mode``,
        ``use_miss`` as for ``prepare_matmul_cuda``. ``grid``: workgroups of the
        persistent launch (default: as many as are resident at once, at most the
        largest phase)."""
        if not self._entered:
            raise RuntimeError("HipRuntime must be entered before preparing a batched sweep")
        return HipBatchedSweep(self, entries, persistent=persistent, grid=grid)

    def _prepare_matmul_cuda(self, grg: BoundGRG, spec: _CudaMatmulSpec) -> _HipPreparedMatmul:
        if spec.direction == Direction.UP and self.layout.pair.plan_up is None:
            raise ValueError("UP plan is not configured")
        if spec.direction == Direction.DOWN and self.layout.pair.plan_down is None:
            raise ValueError("DOWN plan is not configured")
        return _HipPreparedMatmul(self, self._artifacts[int(grg._artifact_index)], spec)

    # ---- artifact construction ------------------------------------------

    def _to_dev(self, arr: np.ndarray, dtype=None):
        torch = _torch()
        a = np.ascontiguousarray(arr)
        if a.size == 0:
            a = np.zeros(1, dtype=a.dtype)
        t = torch.from_numpy(a).to(device=self.device)
        return t if dtype is None else t.to(dtype)

    def _pinned(self, values: np.ndarray):
        torch = _torch()
        host = torch.empty((max(1, int(np.asarray(values).size)),), dtype=torch.int32, pin_memory=True)
        if np.asarray(values).size:
            copy_struct_checked(host.numpy()[: np.asarray(values).size], np.asarray(values), label="streamed block structure")
        return host

    def _merge_parts(self, N: int, parts):
        """One CSR over all N rows from parts (row_ids int64[rows], indptr int64[rows + 1], cols int32[nnz]).

        ``row_ids`` are the (new) global rows of a part's local rows; entries keep
        their order within a row and parts are laid out in list order within each
        row. Returns (ptr int32[N + 1], idx int32[max(E, 1)], degree int64[N]) on device.
        """
        torch = _torch()
        dev = self.device
        deg = torch.zeros(N, dtype=torch.int64, device=dev)
        for rows, ptr, _ in parts:
            deg.index_add_(0, rows, ptr[1:] - ptr[:-1])
        ptr_all = torch.zeros(N + 1, dtype=torch.int64, device=dev)
        torch.cumsum(deg, 0, out=ptr_all[1:])
        total = int(ptr_all[-1].item())
        if total > np.iinfo(np.int32).max:
            raise NotImplementedError("more than 2^31 - 1 resident entries in one sweep orientation")
        idx = torch.zeros(max(total, 1), dtype=torch.int32, device=dev)
        cursor = ptr_all[:-1].clone()
        for rows, ptr, cols in parts:
            d = ptr[1:] - ptr[:-1]
            nnz = int(cols.numel())
            if nnz == 0:
                continue
            local = torch.repeat_interleave(torch.arange(rows.numel(), device=dev), d)
            pos = cursor[rows[local]] + (torch.arange(nnz, device=dev) - ptr[:-1][local])
            idx[pos] = cols
            cursor.index_add_(0, rows, d)
        return ptr_all.to(torch.int32), idx, deg

    def _build_artifact(self, ai: int, al, st) -> _HipArtifact:
        from pygrgl_spmv.backends.hip.host import build_tiles, cluster_order, locality_order, merged_keys

        torch = _torch()
        lay = self.layout
        dev = self.device
        off = np.asarray(st.level_offsets, dtype=np.int64)
        L = off.size - 1
        S, M, N, I = st.num_samples, st.num_mutations, st.num_nodes, st.num_individuals
        art = _HipArtifact(path=al.path, state=st, level_offsets=off)
        D = art.dev
        up_plan = {(b.dst_level, b.src_level): b for b in al.blocks_up}
        dn_plan = {(b.dst_level, b.src_level): b for b in al.blocks_down}
        host_blocks = []
        for blk in iter_artifact_blocks(al.path):
            if blk.nnz == 0:
                continue
            idx = np.asarray(blk.indices)
            if idx.dtype != np.int32:
                idx = _i32(idx, f"A_blocks[{blk.dst_level}][{blk.src_level}].indices")
            host_blocks.append((blk.dst_level, blk.src_level, np.asarray(blk.indptr, dtype=np.int64), idx))
        # Device-private node order: a permutation inside every level that keeps nodes
        # sharing descendants adjacent (samples stay the prefix of level 0).
        new_pos = locality_order(N, S, off, host_blocks, degree_window=self.degree_window,
                                 degree_min_mean=self.degree_min_mean)
        if self.cluster_window > 1:
            new_pos = cluster_order(new_pos, off, host_blocks, window=self.cluster_window,
                                    degree_window=self.cluster_degree_window, min_mean=self.degree_min_mean,
                                    min_rows=self.cluster_min_rows)
        new_pos_t = self._to_dev(new_pos)
        up_parts, dn_parts = [], []
        up_stream = [[] for _ in range(L)]
        dn_stream = [[] for _ in range(L)]
        with torch.cuda.device(dev):
            for d, s, ptr_h, idx_h in host_blocks:
                rows_d = int(off[d + 1] - off[d])
                rows_s = int(off[s + 1] - off[s])
                bp = up_plan.get((d, s))
                if bp is not None:
                    if bp.resident:
                        ptr_d = self._to_dev(ptr_h)
                        cols = new_pos_t[self._to_dev(idx_h).to(torch.int64) + int(off[s])].to(torch.int32)
                        up_parts.append((d, s, (new_pos_t[int(off[d]) : int(off[d + 1])], ptr_d, cols)))
                    else:
                        rp, ri = _remapped_block(ptr_h, idx_h, new_pos, off, d, s, transpose=False)
                        up_stream[d].append(_StreamStep(
                            dst_level=d, src_level=s, slot=int(bp.slot), host_indices=self._pinned(ri),
                            host_indptr=self._pinned(rp), row_lo=int(off[d]), row_hi=int(off[d + 1]),
                            src_offset=int(off[s]), nnz=int(ri.size)))
                bp = dn_plan.get((d, s))
                if bp is not None:
                    if bp.resident:
                        ptr_d = self._to_dev(ptr_h)
                        cols_d = self._to_dev(idx_h).to(torch.int64)
                        row_of = torch.repeat_interleave(torch.arange(rows_d, device=dev), ptr_d[1:] - ptr_d[:-1])
                        keys, order = torch.sort(cols_d, stable=True)
                        parents = new_pos_t[row_of[order] + int(off[d])].to(torch.int32)
                        counts = torch.bincount(keys, minlength=rows_s)
                        tptr = torch.zeros(rows_s + 1, dtype=torch.int64, device=dev)
                        torch.cumsum(counts, 0, out=tptr[1:])
                        dn_parts.append((d, s, (new_pos_t[int(off[s]) : int(off[s + 1])], tptr, parents)))
                    else:
                        rp, ri = _remapped_block(ptr_h, idx_h, new_pos, off, d, s, transpose=True)
                        dn_stream[s].append(_StreamStep(
                            dst_level=s, src_level=d, slot=int(bp.slot), host_indices=self._pinned(ri),
                            host_indptr=self._pinned(rp), row_lo=int(off[s]), row_hi=int(off[s + 1]),
                            src_offset=int(off[d]), nnz=int(ri.size)))
            del host_blocks
            deg_up = deg_dn = None
            if lay.pair.plan_up is not None:
                parts = [p for _, _, p in sorted(up_parts, key=lambda t: (t[0], t[1]))]
                D["up_ptr"], D["up_idx"], deg_up = self._merge_parts(N, parts)
            if lay.pair.plan_down is not None:
                parts = [p for _, _, p in sorted(dn_parts, key=lambda t: (t[1], -t[0]))]
                D["dn_ptr"], D["dn_idx"], deg_dn = self._merge_parts(N, parts)
            del up_parts, dn_parts
            # selectors, in the new node ids
            sel_ptr = np.asarray(st.sel_mut.indptr, dtype=np.int64)
            sel_idx = new_pos[np.asarray(st.sel_mut.indices, dtype=np.int64)]
            miss_ptr = np.asarray(st.sel_miss.indptr, dtype=np.int64)
            miss_idx = new_pos[np.asarray(st.sel_miss.indices, dtype=np.int64)]
            req = lay.requirements
            if lay.pair.plan_up is not None:
                D["sel_ptr"], D["sel_idx"] = self._to_dev(_i32(sel_ptr, "sel_mut.indptr")), self._to_dev(_i32(sel_idx, "sel_mut.indices"))
                if req.need_up_miss_output:
                    D["miss_ptr"], D["miss_idx"] = self._to_dev(_i32(miss_ptr, "sel_miss.indptr")), self._to_dev(_i32(miss_idx, "sel_miss.indices"))
            deg_mut_t = deg_miss_t = None
            sel_t_host = miss_t_host = None
            if lay.pair.plan_down is not None:
                tp, ti = _csr_transpose(sel_ptr, sel_idx, M, N)
                sel_t_host = (_i32(tp, "sel_mut^T.indptr"), _i32(ti, "sel_mut^T.indices"), M)
                D["selT_ptr"], D["selT_idx"] = self._to_dev(sel_t_host[0]), self._to_dev(sel_t_host[1])
                deg_mut_t = np.diff(tp)
                if req.need_down_miss_input:
                    tp, ti = _csr_transpose(miss_ptr, miss_idx, M, N)
                    miss_t_host = (_i32(tp, "sel_miss^T.indptr"), _i32(ti, "sel_miss^T.indices"), M)
                    D["missT_ptr"], D["missT_idx"] = self._to_dev(miss_t_host[0]), self._to_dev(miss_t_host[1])
                    deg_miss_t = np.diff(tp)
            # node / sample maps through the new order
            node_perm = np.empty(N, dtype=np.int64)
            node_perm[new_pos] = np.asarray(st.node_perm, dtype=np.int64)          # new internal -> original id
            inv_perm = new_pos[np.asarray(st.inv_node_perm, dtype=np.int64)]        # original id -> new internal
            D["node_perm"] = self._to_dev(_i32(node_perm, "node_perm"))
            D["inv_perm"] = self._to_dev(_i32(inv_perm, "inv_node_perm"))
            s2i = np.asarray(st.sample_to_individual, dtype=np.int64)
            sample_at = np.empty(S, dtype=np.int64)
            sample_at[new_pos[:S]] = np.arange(S)                                    # new row -> sample id
            D["in_map"] = self._to_dev(_i32(sample_at, "input map"))
            D["in_map_ind"] = self._to_dev(_i32(s2i[sample_at], "input map (individuals)"))
            D["out_map"] = self._to_dev(_i32(new_pos[:S], "output map"))
            order = np.argsort(s2i, kind="stable")
            ind_ptr = np.zeros(I + 1, dtype=np.int64)
            np.cumsum(np.bincount(s2i, minlength=I), out=ind_ptr[1:])
            D["ind_ptr"], D["ind_idx"] = self._to_dev(_i32(ind_ptr, "individual.indptr")), self._to_dev(_i32(new_pos[order], "individual.indices"))
            npdt = np.dtype(lay.dtype)
            if req.need_init_xtx and st.coalescence_counts is not None:
                coal2 = np.empty(N, dtype=np.float64)
                coal2[new_pos] = 2.0 * np.asarray(st.coalescence_counts, dtype=np.float64)
                D["coal2"] = self._to_dev(coal2.astype(npdt))
            if req.need_init_vector:
                if lay.pair.plan_up is not None:
                    D["vec_up"] = self._to_dev(np.asarray(st.init_vector_up_bias, dtype=npdt))
                if lay.pair.plan_down is not None:
                    vd = np.asarray(st.init_vector_down_bias, dtype=np.float64)
                    D["vec_dn"] = self._to_dev(vd.astype(npdt))
                    D["vec_dn_ind"] = self._to_dev(np.bincount(s2i, weights=vd, minlength=I).astype(npdt))
            if req.need_init_xtx and st.init_xtx_up_bias is not None:
                if lay.pair.plan_up is not None:
                    D["xtx_up"] = self._to_dev(np.asarray(st.init_xtx_up_bias, dtype=npdt))
                if lay.pair.plan_down is not None:
                    xd = np.asarray(st.init_xtx_down_bias, dtype=np.float64)
                    D["xtx_dn"] = self._to_dev(xd.astype(npdt))
                    D["xtx_dn_ind"] = self._to_dev(np.bincount(s2i, weights=xd, minlength=I).astype(npdt))
            # row-batch tables and heavy rows for every sweep configuration
            configs = {}
            if deg_up is not None:
                configs["up"] = deg_up.cpu().numpy()
            if deg_dn is not None:
                base = deg_dn.cpu().numpy() + deg_mut_t
                configs["down"] = base
                if deg_miss_t is not None:
                    configs["down_miss"] = base + deg_miss_t
            tables = []
            cursor = 0
            for name, deg in configs.items():
                for l in range(L):
                    sched, h = level_schedule(deg, int(off[l]), int(off[l + 1]), heavy_min=int(lay.heavy_threshold), cursor=cursor,
                                              split_min=self.split_min, split_chunk=self.split_chunk)
                    art.sched[(name, l)] = sched
                    art.max_items = max(art.max_items, sched.num_items)
                    tables.append(h)
                    cursor += h.size
            cap = lay.schedule_capacity[ai] if ai < len(lay.schedule_capacity) else cursor
            if cursor > cap:
                raise RuntimeError(f"internal error: {cursor} schedule entries exceed the planned capacity {cap}")
            D["sched"] = self._to_dev(np.concatenate(tables).astype(np.int32) if tables else np.zeros(1, np.int32))
            if lay.tiles and (self.use_tiles or self.use_groups):
                # tile tables of the LDS-staged level launches and row groups of the row-group
                # launches, one set per sweep configuration
                seg_host = {}
                if "up_ptr" in D:
                    seg_host["up"] = [(D["up_ptr"].cpu().numpy(), D["up_idx"].cpu().numpy(), N)]
                if "dn_ptr" in D:
                    dn = (D["dn_ptr"].cpu().numpy(), D["dn_idx"].cpu().numpy(), N)
                    seg_host["down"] = [dn, sel_t_host]
                    if miss_t_host is not None:
                        seg_host["down_miss"] = [dn, sel_t_host, miss_t_host]
                for name, segs_h in seg_host.items():
                    if name not in configs:
                        continue
                    heavy = [art.sched[(name, l)].heavy_threshold for l in range(L)]
                    if self.use_tiles:
                        t = build_tiles(off, segs_h, heavy, max_staged=self.tile_max_staged, max_rows=self.tile_max_rows,
                                        max_entries=self.tile_max_entries)
                        D[f"{name}_tile_row"] = self._to_dev(t.tile_row)
                        D[f"{name}_tile_src"] = self._to_dev(t.tile_src)
                        D[f"{name}_src_list"] = self._to_dev(t.src_list.view(np.int32))
                        D[f"{name}_lidx"] = self._to_dev(t.lidx.view(np.int16))
                        art.tiles[name] = (t.level_tile_off.copy(), t.level_max_staged.copy(), t.level_max_rows.copy(),
                                           t.level_max_entries.copy())
                        eptr_h = t.eptr
                        del t
                    else:
                        # merged row pointers: each row's entries over all segments, in segment order
                        deg = sum(np.diff(np.asarray(p_, dtype=np.int64)) for p_, _, _ in segs_h)
                        eptr64 = np.zeros(N + 1, dtype=np.int64)
                        np.cumsum(deg, out=eptr64[1:])
                        eptr_h = _i32(eptr64, f"{name} merged row pointers")
                    D[f"{name}_eptr"] = self._to_dev(eptr_h)
                    keys = merged_keys(eptr_h, segs_h) if len(segs_h) > 1 else segs_h[0][1]
                    if len(segs_h) > 1:
                        D[f"{name}_ekey"] = self._to_dev(keys.view(np.int32))
                    if self.use_panels and self.use_tiles:
                        self._build_panels(art, name, eptr_h, keys, off)
                    if self.use_groups and self._dtype_code == ffi.F64:
                        self._build_groups(art, name, eptr_h, keys, off)
                    del keys, eptr_h
                del seg_host
            # streamed ops with slot-reuse links, and their events
            art.up_stream = relink_stream_dependencies(up_stream, direction=Direction.UP)
            art.down_stream = relink_stream_dependencies(dn_stream, direction=Direction.DOWN)
            art.up_events = [[(torch.cuda.Event(), torch.cuda.Event()) for _ in ops] for ops in art.up_stream]
            art.down_events = [[(torch.cuda.Event(), torch.cuda.Event()) for _ in ops] for ops in art.down_stream]
        return art

    # ---- launch construction ----------------------------------------------

    def _tune_context_fields(self, al, art: _HipArtifact) -> dict:
        """What the tuned shapes of one artifact depend on besides the launch itself."""
        torch = _torch()
        lib = Path(ffi.LIB_PATH).stat()
        props = torch.cuda.get_device_properties(self.device)
        resident = "".join("1" if b.resident else "0" for b in (*al.blocks_up, *al.blocks_down))
        return {"artifact": tunecache.file_digest(art.path), "dtype": int(self._dtype_code), "abi": ffi.ABI_VERSION,
                "library": [int(lib.st_size), int(lib.st_mtime_ns)],
                # the architecture and CU count, not the marketing name (some tools report it empty)
                "device": [str(getattr(props, "gcnArchName", "")), int(props.multi_processor_count)],
                "split": [self.split_min, self.split_chunk], "heavy": int(self.layout.heavy_threshold),
                "tiles": [bool(self.layout.tiles and self.use_tiles), self.tile_max_staged, self.tile_max_rows,
                          self.tile_max_entries],
                "panels": [self.use_panels, self.panel_min_mean, self.panel_min_rows, self.panel_rows, self.panel_staged],
                "order": [self.degree_window, self.degree_min_mean, self.cluster_window, self.cluster_degree_window,
                          self.cluster_min_rows],
                "resident": hashlib.blake2b(resident.encode(), digest_size=8).hexdigest()}

    def _flush_tune_cache(self) -> None:
        for ctx in self._tune_ctx.values():
            ctx.flush()

    def _shape_admissible(self, a: ffi.GrgHipGatherArgs, shape, lds) -> bool:
        """Whether a cached shape is one the kernel accepts for this launch."""
        cpl, kl, es, unroll, rows, fl = shape
        if (cpl, kl) not in column_shapes(a.dtype, a.k, lds) and not (fl & (ffi.FLAG_TILED | ffi.FLAG_PANEL | ffi.FLAG_GROUPS)):
            return False
        if fl & ffi.FLAG_GROUPS and (not a.grp_off or a.dtype != ffi.F64 or kl * cpl != ffi.GROUP_ROWS):
            return False
        if cpl < 1 or a.k % cpl or any(int(ld) % cpl for ld in lds) or kl < 1 or kl & (kl - 1) or es < 1 or es & (es - 1):
            return False
        if (fl & ffi.FLAG_TILED and not a.tile_row) or (fl & ffi.FLAG_PANEL and not a.pan_pass):
            return False
        if fl & ffi.FLAG_MERGED and not a.ekey:
            return False
        if fl & ffi.FLAG_PANEL:
            return es == 1 and unroll in (4, 8) and rows == panel_rows_per_group(a.panel_rows, kl) != 0
        return (unroll, rows) in ((4, 1), (8, 1), (16, 1), (4, 2), (8, 2), (4, 4))

    def _elem(self) -> int:
        return 8 if self._dtype_code == ffi.F64 else 4

    def _build_panels(self, art: _HipArtifact, cfg: str, eptr: np.ndarray, keys: np.ndarray, off: np.ndarray) -> None:
        """Panels of the level launches of sweep configuration ``cfg`` whose light rows
        average at least ``panel_min_mean`` entries (``grg_host_build_panels``)."""
        from pygrgl_spmv.backends.hip.host import build_panels

        D = art.dev
        for l in range(off.size - 1):
            lo, hi = int(off[l]), int(off[l + 1])
            sched = art.sched.get((cfg, l))
            if sched is None or hi - lo < max(1, self.panel_min_rows) or sched.mean_degree < self.panel_min_mean:
                continue
            pn = build_panels(eptr, keys, lo, hi, panel_rows=self.panel_rows, max_staged=self.panel_staged,
                              max_pass_entries=PANEL_PASS_ENTRIES, heavy_threshold=int(sched.heavy_threshold))
            if pn.max_pass_entries + 32 > PANEL_MAX_PASS_BYTES or pn.max_passes > PANEL_MAX_PASSES:
                continue
            tag = f"{cfg}_pan{l}"
            D[f"{tag}_pass"] = self._to_dev(pn.pan_pass)
            D[f"{tag}_src_off"] = self._to_dev(pn.pass_src)
            D[f"{tag}_ent_off"] = self._to_dev(pn.pass_ent)
            D[f"{tag}_rowoff"] = self._to_dev(pn.pass_rowoff.view(np.int16))
            D[f"{tag}_src"] = self._to_dev(pn.src_list.view(np.int32))
            D[f"{tag}_lidx"] = self._to_dev(pn.lidx.view(np.int8))
            art.panels[(cfg, l)] = (pn.num_panels, pn.max_pass_entries, pn.max_passes)

    def _build_groups(self, art: _HipArtifact, cfg: str, eptr: np.ndarray, keys: np.ndarray, off: np.ndarray) -> None:
        """Row groups (``grg_host_build_groups``) of the level launches of sweep configuration
        ``cfg`` whose light rows average at least ``group_min_mean`` entries, kept where
        neighbouring rows share sources (fewer group entries than ``group_max_fill`` of the
        rows' entries)."""
        from pygrgl_spmv.backends.hip.host import build_groups

        D = art.dev
        for l in range(off.size - 1):
            lo, hi = int(off[l]), int(off[l + 1])
            sched = art.sched.get((cfg, l))
            if sched is None or (cfg == "up" and l == 0) or hi - lo < max(1, self.group_min_rows):
                continue
            if sched.mean_degree < self.group_min_mean:
                continue
            gr = build_groups(eptr, keys, lo, hi, heavy_threshold=int(sched.heavy_threshold))
            if gr.num_entries > self.group_max_fill * max(1, int(eptr[hi]) - int(eptr[lo])):
                continue
            tag = f"{cfg}_grp{l}"
            D[f"{tag}_off"] = self._to_dev(gr.off)
            D[f"{tag}_store"] = self._to_dev(gr.store.view(np.int16))
            D[f"{tag}_key"] = self._to_dev(gr.key.view(np.int32))
            D[f"{tag}_mask"] = self._to_dev(gr.mask.view(np.int16))
            art.groups[(cfg, l)] = (gr.num_groups, gr.num_entries)

    def _group_fields(self, art: _HipArtifact, cfg: str, level: int):
        """Row-group fields of one level launch of sweep configuration ``cfg``, or None."""
        meta = art.groups.get((cfg, level))
        if meta is None:
            return None
        D = art.dev
        tag = f"{cfg}_grp{level}"
        return {"grp_off": D[f"{tag}_off"].data_ptr(), "grp_store": D[f"{tag}_store"].data_ptr(),
                "grp_key": D[f"{tag}_key"].data_ptr(), "grp_mask": D[f"{tag}_mask"].data_ptr(), "num_groups": int(meta[0])}

    def _panel_fields(self, art: _HipArtifact, cfg: str, level: int):
        """Panel fields of one level launch of sweep configuration ``cfg``, or None."""
        meta = art.panels.get((cfg, level))
        if meta is None:
            return None
        D = art.dev
        tag = f"{cfg}_pan{level}"
        return {"pan_pass": D[f"{tag}_pass"].data_ptr(), "pass_src": D[f"{tag}_src_off"].data_ptr(),
                "pass_ent": D[f"{tag}_ent_off"].data_ptr(), "pass_rowoff": D[f"{tag}_rowoff"].data_ptr(),
                "pan_src": D[f"{tag}_src"].data_ptr(), "pan_lidx": D[f"{tag}_lidx"].data_ptr(), "num_panels": int(meta[0]),
                "panel_rows": self.panel_rows, "panel_staged": self.panel_staged, "panel_max_entries": int(meta[1]),
                "panel_max_passes": int(meta[2])}

    def _tile_fields(self, art: _HipArtifact, cfg: str, level: int):
        """Tile fields of one level launch of sweep configuration ``cfg``, or None."""
        tl = art.tiles.get(cfg)
        if tl is None:
            return None
        level_tile_off, level_max_staged, level_max_rows, level_max_entries = tl
        D = art.dev
        t0, t1 = int(level_tile_off[level]), int(level_tile_off[level + 1])
        if t1 <= t0:
            return None
        return {"tile_row": D[f"{cfg}_tile_row"].data_ptr() + 4 * t0, "tile_src": D[f"{cfg}_tile_src"].data_ptr() + 4 * t0,
                "src_list": D[f"{cfg}_src_list"].data_ptr(), "eptr": D[f"{cfg}_eptr"].data_ptr(),
                "lidx": D[f"{cfg}_lidx"].data_ptr(), "num_tiles": t1 - t0, "tile_lds_rows": int(level_max_staged[level]),
                "tile_max_rows": int(level_max_rows[level]), "tile_max_entries": int(level_max_entries[level]),
                "ekey": D[f"{cfg}_ekey"].data_ptr() if f"{cfg}_ekey" in D else 0}

    def _args(self, *, rows, k, out, ld_out, segs=(), maps=(), row_scale=None, col_vec=None, accumulate=False,
              sched: _LevelSchedule | None = None, sched_base: int = 0, mean_degree: float = 1.0,
              out_base=0, tiles: dict | None = None, panels: dict | None = None, groups: dict | None = None,
              partials=None) -> ffi.GrgHipGatherArgs:
        """Launch descriptor for rows [rows); ``sched`` adds the level's heavy rows,
        ``tiles`` (from ``_tile_fields``) the level's tile tables, ``panels`` (from
        ``_panel_fields``) its panels, ``groups`` (from ``_group_fields``) its row groups."""
        a = ffi.GrgHipGatherArgs()
        a.dtype = self._dtype_code
        a.accumulate = 1 if accumulate else 0
        a.row_begin, a.row_end = int(rows[0]), int(rows[1])
        a.k = int(k)
        if sched is not None:
            mean_degree = sched.mean_degree
            a.heavy_threshold = sched.heavy_threshold
            a.heavy_rows, a.num_heavy = sched_base + 4 * sched.heavy_offset, sched.num_heavy
            if sched.num_items:
                a.split_ptr = sched_base + 4 * sched.split_ptr_offset
                a.split_row = sched_base + 4 * sched.split_row_offset
                a.split_items, a.split_chunk = sched.num_items, sched.split_chunk
                a.partials = (self._partials if partials is None else partials).data_ptr()
        else:
            a.heavy_threshold, a.heavy_rows, a.num_heavy = -1, 0, 0
        a.flags = 0
        a.unroll = KERNEL_UNROLL
        a.rows_per_group = 1
        lds = [ld_out] + [s[3] for s in segs] + [m[2] for m in maps]
        a.cpl, a.kl, a.es = launch_shape(self._dtype_code, int(k), lds, mean_degree)
        if tiles:
            for name, value in tiles.items():
                setattr(a, name, int(value))
        if panels:
            for name, value in panels.items():
                setattr(a, name, int(value))
        if groups:
            for name, value in groups.items():
                setattr(a, name, int(value))
        for name, value in self.fixed_shape.items():
            setattr(a, name, int(value))
        if (a.flags & ffi.FLAG_MERGED) and not a.ekey:
            a.flags &= ~ffi.FLAG_MERGED
        if a.flags & ffi.FLAG_STREAM:
            # the streamed path reads merged keys, or the indices of a single segment
            if a.nseg > 1 and a.ekey:
                a.flags |= ffi.FLAG_MERGED
            if a.nseg < 1 or (a.nseg > 1 and not a.ekey) or (a.flags & (ffi.FLAG_TILED | ffi.FLAG_PARTIALS)):
                a.flags &= ~ffi.FLAG_STREAM
            else:
                a.unroll = 4 if a.unroll == 4 else 8
                a.rows_per_group = 1
        if a.flags & ffi.FLAG_PANEL:
            # a panel launch needs the level's panels and lane groups that own 1, 2 or 4 rows
            if not panels:
                a.flags &= ~ffi.FLAG_PANEL
            else:
                a.flags &= ~(ffi.FLAG_TILED | ffi.FLAG_STREAM | ffi.FLAG_MERGED | ffi.FLAG_GROUP_KEYS)
                while a.kl > 1 and not panel_rows_per_group(a.panel_rows, a.kl):
                    a.kl //= 2
                a.es = 1
                a.rows_per_group = panel_rows_per_group(a.panel_rows, a.kl)
                if not a.rows_per_group:
                    a.flags &= ~ffi.FLAG_PANEL
                    a.rows_per_group = 1
                elif a.unroll not in (4, 8):
                    a.unroll = 8
        if a.flags & ffi.FLAG_GROUPS:
            # a row-group launch needs the level's groups (FP64) and one 16-column chunk
            if not groups or a.dtype != ffi.F64 or a.cpl not in (1, 2, 4):
                a.flags &= ~ffi.FLAG_GROUPS
            else:
                a.flags &= ~(ffi.FLAG_TILED | ffi.FLAG_PANEL | ffi.FLAG_STREAM | ffi.FLAG_MERGED | ffi.FLAG_GROUP_KEYS)
                a.kl, a.es, a.rows_per_group = ffi.GROUP_ROWS // a.cpl, 1, 1
                if a.unroll not in (4, 8, 16):
                    a.unroll = 8
        if a.flags & ffi.FLAG_TILED:
            if not tiles:
                a.flags &= ~ffi.FLAG_TILED
            else:
                while a.kl > 1 and a.kl * a.cpl * self._elem() > ffi.TILE_CHUNK_BYTES:
                    a.kl //= 2
        if a.flags & (ffi.FLAG_XCD_REMAP | ffi.FLAG_STREAM):
            a.flags &= ~(ffi.FLAG_XCD_RUNS | XCD_RUN_BITS)  # runs apply to the light path, tiles and panels
        if a.kl * a.es > 64 and not a.flags & (ffi.FLAG_PANEL | ffi.FLAG_GROUPS):
            a.es = 64 // a.kl
        if (a.flags & ffi.FLAG_GROUP_KEYS) and ((not a.flags & ffi.FLAG_MERGED and a.nseg != 1) or a.rows_per_group > 1
                                                or a.flags & (ffi.FLAG_TILED | ffi.FLAG_PANEL | ffi.FLAG_STREAM)):
            a.flags &= ~ffi.FLAG_GROUP_KEYS  # group-loaded entries need one row per lane group, merged or one segment
        a.out, a.ld_out, a.out_base = int(out), int(ld_out), int(out_base)
        a.nseg = len(segs)
        for i, (ptr, idx, src, ld, base) in enumerate(segs):
            a.seg[i].ptr, a.seg[i].idx, a.seg[i].src, a.seg[i].ld, a.seg[i].ptr_base = int(ptr), int(idx), int(src), int(ld), int(base)
        a.nmap = len(maps)
        for i, (mp, src, ld, lo, hi) in enumerate(maps):
            a.map[i].map, a.map[i].src, a.map[i].ld, a.map[i].row_lo, a.map[i].row_hi = int(mp or 0), int(src), int(ld), int(lo), int(hi)
        if row_scale is not None:
            a.row_scale, a.scale_base = int(row_scale[0]), int(row_scale[1])
        a.col_vec = int(col_vec or 0)
        return a

    def _time_launch(self, args: ffi.GrgHipGatherArgs, reps: int = 3) -> float:
        """Best-of-``reps`` device time (ms) of one launch on the root stream, after a warm launch."""
        torch = _torch()
        root = self._root_stream
        ptr = int(root.cuda_stream)
        self._lib.gather(args, ptr)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
        ev[0].record(root)
        for i in range(reps):
            self._lib.gather(args, ptr)
            ev[i + 1].record(root)
        ev[-1].synchronize()
        return min(ev[i].elapsed_time(ev[i + 1]) for i in range(reps))

    def _tune(self, a: ffi.GrgHipGatherArgs, key, private_out: bool = False) -> None:
        """Per-level autotuner, run when an op is prepared (never in ``op()``; the
        timing launches write only the level's node rows, or, with ``private_out``
        (write-out launches, whose rows are the caller-visible arena), a temporary
        buffer of the same shape). Stage 1 times every
        admissible (cpl, kl, es) shape with 8 gathers in flight and one row per lane
        group; the later stages try, each on the best launch so far, the other
        (gathers in flight, rows per group) pairs, merged entries, entries loaded by
        the whole lane group, the XCD remap and XCD runs, the streamed form, panels
        and the tiled form. The fastest launch is cached per artifact, sweep
        configuration, level and k."""
        torch = _torch()
        lds = [a.ld_out] + [a.seg[i].ld for i in range(a.nseg)] + [a.map[i].ld for i in range(a.nmap)]
        # the leading dimensions decide which columns-per-lane widths are admissible, so a
        # shape tuned for one set of buffers (the runtime's arena at ld = max_k, a batched
        # sweep's private buffers at ld = k) is only reused for launches with the same ones
        key = tuple(key) + (tuple(int(v) for v in lds),)
        best = self._tune_cache.get(key)
        ctx = self._tune_ctx.get(key[0]) if self.autotune else None
        disk_key = tunecache.launch_key(key[1], key[2], key[3], lds) if ctx is not None else None
        if best is None and ctx is not None and a.row_end > a.row_begin:
            cached = ctx.get(disk_key)
            if cached is not None and self._shape_admissible(a, cached, lds):
                best = cached
                self._tune_cache[key] = best
                self.tune_log[key] = {"rows": int(a.row_end - a.row_begin), "best": (None,) + best, "cached": True}
        if best is None and self.autotune and a.row_end > a.row_begin:
            trial = ffi.GrgHipGatherArgs()
            ctypes.pointer(trial)[0] = a
            elem = 8 if a.dtype == ffi.F64 else 4
            timings = []
            scratch = None
            if private_out:
                scratch = torch.empty(((a.row_end - a.out_base) * a.ld_out,), dtype=self._torch_dtype(), device=self.device)
                trial.out = scratch.data_ptr()

            def time_shape(cpl, kl, es, unroll, rows, fl):
                trial.cpl, trial.kl, trial.es, trial.unroll, trial.rows_per_group, trial.flags = cpl, kl, es, unroll, rows, fl
                timings.append((self._time_launch(trial), cpl, kl, es, unroll, rows, fl))

            with torch.cuda.device(self.device):
                for cpl, kl in column_shapes(a.dtype, a.k, lds):
                    for es in (1, 2, 4, 8, 16, 32, 64):
                        if kl * es <= 64:
                            time_shape(cpl, kl, es, 8, 1, 0)
                _, cpl, kl, es, _, _, _ = min(timings)
                for unroll, rows in ((4, 1), (16, 1), (4, 2), (8, 2), (4, 4)):
                    if unroll * rows * cpl * elem > TUNE_INFLIGHT_BYTES:  # loads in flight per lane, bounded
                        continue
                    time_shape(cpl, kl, es, unroll, rows, 0)
                _, cpl, kl, es, unroll, rows, _ = min(timings)
                if a.ekey:
                    # the merged-entry light path on the best shape and its neighbours in es
                    for es_m in sorted({es, max(1, es // 2), min(64 // kl, es * 2)}):
                        time_shape(cpl, kl, es_m, unroll, rows, ffi.FLAG_MERGED)
                    _, cpl, kl, es, unroll, rows, fl = min(timings)
                    # one merged row per lane group, its keys loaded by the whole group
                    for es_g in sorted({es, 64 // kl}):
                        time_shape(cpl, kl, es_g, unroll, 1, ffi.FLAG_MERGED | ffi.FLAG_GROUP_KEYS)
                elif a.nseg == 1:
                    # one row per lane group, its indices loaded by the whole group
                    time_shape(cpl, kl, es, unroll, 1, ffi.FLAG_GROUP_KEYS)
                # block orders over the XCDs (the remap, and runs of consecutive blocks with all
                # XCDs in one window), on the best launch so far and on the best with
                # group-loaded entries: the orders change which of the two wins
                starts = [min(timings)]
                grouped = [t for t in timings if t[6] & ffi.FLAG_GROUP_KEYS]
                if grouped and min(grouped)[1:] != starts[0][1:]:
                    starts.append(min(grouped))
                for _, cpl, kl, es, unroll, rows, fl in starts:
                    keep = fl & (ffi.FLAG_MERGED | ffi.FLAG_GROUP_KEYS)
                    time_shape(cpl, kl, es, unroll, rows, keep | ffi.FLAG_XCD_REMAP)
                    for log_run in XCD_RUN_LOGS:
                        time_shape(cpl, kl, es, unroll, rows, keep | xcd_runs_flags(log_run))
                if a.nseg == 1 or a.ekey:
                    # the streamed (software-pipelined) light path on the best column shape
                    _, cpl, kl, es, unroll, rows, fl = min(timings)
                    base = ffi.FLAG_STREAM | (ffi.FLAG_MERGED if a.nseg > 1 else 0)
                    for es_s in sorted({es, max(1, es // 2), min(64 // kl, es * 2)}):
                        for u in (4, 8):
                            time_shape(cpl, kl, es_s, u, 1, base)
                if a.pan_pass:
                    # the panel form: lane groups of kl lanes owning panel_rows * kl / 1024 rows each
                    n_before = len(timings)
                    for cpl, kl in column_shapes(a.dtype, a.k, lds):
                        for klp in sorted({kl, max(1, kl // 2)}):
                            rpg = panel_rows_per_group(a.panel_rows, klp)
                            if not rpg or a.panel_staged * klp > 2 * PANEL_BLOCK or panel_lds_bytes(
                                    a.dtype, cpl, klp, a.panel_staged, a.panel_rows, a.panel_max_entries,
                                    a.panel_max_passes) > TILE_LDS_LIMIT:
                                continue
                            for u in (4, 8):  # two or three passes in flight
                                time_shape(cpl, klp, 1, u, rpg, ffi.FLAG_PANEL)
                    if len(timings) > n_before:
                        _, cpl, kl, es, unroll, rows, _ = min(timings[n_before:])
                        time_shape(cpl, kl, es, unroll, rows, ffi.FLAG_PANEL | ffi.FLAG_XCD_REMAP)
                        for log_run in XCD_RUN_LOGS:
                            time_shape(cpl, kl, es, unroll, rows, ffi.FLAG_PANEL | xcd_runs_flags(log_run))
                if a.tile_row:
                    # the LDS-staged form of the same launch: column chunks of at most
                    # TILE_CHUNK_BYTES, every lane-group width
                    n_untiled = len(timings)
                    for cpl, kl in column_shapes(a.dtype, a.k, lds):
                        while kl > 1 and kl * cpl * elem > ffi.TILE_CHUNK_BYTES:
                            kl //= 2
                        if kl * cpl * elem > ffi.TILE_CHUNK_BYTES or tile_lds_bytes(a.dtype, cpl, kl, a.tile_lds_rows, a.tile_max_rows, a.tile_max_entries) > TILE_LDS_LIMIT:
                            continue
                        for es in (1, 2, 4, 8, 16):
                            if kl * es <= 64:
                                time_shape(cpl, kl, es, 8, 1, ffi.FLAG_TILED)
                    if len(timings) > n_untiled:
                        _, cpl, kl, es, unroll, rows, _ = min(timings[n_untiled:])
                        time_shape(cpl, kl, es, unroll, rows, ffi.FLAG_TILED | ffi.FLAG_XCD_REMAP)
                        for log_run in XCD_RUN_LOGS:
                            time_shape(cpl, kl, es, unroll, rows, ffi.FLAG_TILED | xcd_runs_flags(log_run))
                if a.grp_off and a.dtype == ffi.F64:
                    # the row-group form: one wave per 16 consecutive rows, the group's distinct
                    # sources gathered once and summed by MFMA (16-column chunks)
                    n_before = len(timings)
                    # 1 or 2 columns per lane: the 4-column form ties with them on the 500k x 5M
                    # sample level but reads 1.7x as much from memory (10.4 against 6.2 GB)
                    cpls = [c for c, _ in column_shapes(a.dtype, a.k, lds) if c in (1, 2)]
                    for cpl in cpls:
                        for u in (4, 8, 16):
                            time_shape(cpl, ffi.GROUP_ROWS // cpl, 1, u, 1, ffi.FLAG_GROUPS)
                    if len(timings) > n_before:
                        _, cpl, kl, es, unroll, rows, _ = min(timings[n_before:])
                        time_shape(cpl, kl, es, unroll, rows, ffi.FLAG_GROUPS | ffi.FLAG_XCD_REMAP)
                        for log_run in XCD_RUN_LOGS:
                            time_shape(cpl, kl, es, unroll, rows, ffi.FLAG_GROUPS | xcd_runs_flags(log_run))
                top = min(timings)
                finalists = []
                for t in sorted(timings):
                    if all(t[1:] != f[1:] for f in finalists):
                        finalists.append(t)
                    if len(finalists) == TUNE_FINALISTS:
                        break
                if len(finalists) > 1 and top[0] >= TUNE_RETIME_MIN_MS:
                    # near-ties are decided by noise: re-time the fastest few launches,
                    # interleaved, and keep each one's best
                    again = {f[1:]: f[0] for f in finalists}
                    for _ in range(TUNE_RETIME_ROUNDS):
                        for shape in again:
                            trial.cpl, trial.kl, trial.es, trial.unroll, trial.rows_per_group, trial.flags = shape
                            again[shape] = min(again[shape], self._time_launch(trial))
                    winner = min(again, key=again.get)
                    top = (again[winner],) + winner
            best = top[1:]
            self._tune_cache[key] = best
            if ctx is not None:
                ctx.put(disk_key, best)
            tiled = [t for t in timings if t[6] & ffi.FLAG_TILED]
            untiled = [t for t in timings if not t[6] & ffi.FLAG_TILED]
            self.tune_log[key] = {"rows": int(a.row_end - a.row_begin), "best": top,
                                  "untiled_ms": min(untiled)[0] if untiled else None,
                                  "tiled_ms": min(tiled)[0] if tiled else None, "timings": sorted(timings)}
        if best is not None:
            a.cpl, a.kl, a.es, a.unroll, a.rows_per_group, a.flags = best

    def _arena(self) -> _Buffers:
        return _Buffers(state=self._state, ld_state=0, io0=self._io0, io1=self._io1, aux=self._aux, partials=self._partials)

    def _build_steps(self, art: _HipArtifact, spec: _CudaMatmulSpec, bufs: _Buffers | None = None):
        """The prepared sweep: a list of ("batch", array, n) and ("stream", step, args, events) items,
        writing into ``bufs`` (the runtime's arena by default)."""
        bufs = self._arena() if bufs is None else bufs
        D = art.dev
        st = art.state
        off = art.level_offsets
        L = off.size - 1
        k = int(spec.k)
        w = self._elem()
        V = bufs.state.data_ptr()
        ldv = k
        max_k = int(bufs.io0.shape[1])
        io0 = bufs.io0.data_ptr()
        io1 = bufs.io1.data_ptr() if bufs.io1 is not None else 0
        aux = bufs.aux.data_ptr() if bufs.aux is not None else 0
        max_k_aux = int(bufs.aux.shape[1]) if bufs.aux is not None else max_k
        partials = bufs.partials
        sched_base = D["sched"].data_ptr()

        # init terms folded into the sweep (backend_init_mode, as reference grg/__init__.py:498-502)
        maps_init = []
        row_scale_init = None
        col_vec_init = None
        if spec.backend_init_mode == InitMode.MATRIX:
            maps_init.append((D["node_perm"].data_ptr(), io1, max_k, 0, st.num_nodes))
        elif spec.backend_init_mode == InitMode.XTX:
            if "coal2" not in D:
                raise ValueError("init_mode=xtx requires GRG coalescence counts")
            row_scale_init = (D["coal2"].data_ptr(), 0)
        elif spec.backend_init_mode == InitMode.VECTOR:
            col_vec_init = io1  # init vector lives in io1 row 0

        items = []
        batch: list = []

        def flush():
            if batch:
                arr = (ffi.GrgHipGatherArgs * len(batch))(*batch)
                items.append(("batch", arr, len(batch)))
                batch.clear()

        if spec.direction == Direction.UP:
            levels = range(L)
            streams, events, cfg = art.up_stream, art.up_events, "up"
        else:
            levels = range(L - 1, -1, -1)
            streams, events, cfg = art.down_stream, art.down_events, ("down_miss" if spec.use_miss else "down")
        for l in levels:
            lo, hi = int(off[l]), int(off[l + 1])
            segs = []
            maps = list(maps_init)
            if spec.direction == Direction.UP:
                if l > 0:
                    segs.append((D["up_ptr"].data_ptr(), D["up_idx"].data_ptr(), V, ldv, 0))
                else:
                    in_map = D["in_map_ind" if spec.by_individual else "in_map"].data_ptr()
                    maps.insert(0, (in_map, io0, max_k, 0, st.num_samples))
            else:
                segs.append((D["dn_ptr"].data_ptr(), D["dn_idx"].data_ptr(), V, ldv, 0))
                segs.append((D["selT_ptr"].data_ptr(), D["selT_idx"].data_ptr(), io0, max_k, 0))
                if spec.use_miss:
                    segs.append((D["missT_ptr"].data_ptr(), D["missT_idx"].data_ptr(), io1, max_k, 0))
            sched = art.sched[(cfg, l)] if segs else None
            level_args = self._args(rows=(lo, hi), k=k, out=V, ld_out=ldv, segs=segs, maps=maps,
                                    row_scale=row_scale_init, col_vec=col_vec_init, sched=sched, sched_base=sched_base,
                                    tiles=self._tile_fields(art, cfg, l) if segs else None,
                                    panels=self._panel_fields(art, cfg, l) if segs else None,
                                    groups=self._group_fields(art, cfg, l) if segs else None, partials=partials)
            if segs:
                self._tune(level_args, (art.path, cfg, l, k))
            if sched is not None and sched.num_items:
                # partial rows of the level's split rows, same column shape as the level launch
                part = ffi.GrgHipGatherArgs()
                ctypes.pointer(part)[0] = level_args
                part.flags = ffi.FLAG_PARTIALS
                batch.append(part)
            batch.append(level_args)
            for i, step in enumerate(streams[l]):
                flush()
                slot_ptr = self._slot_indptr[step.slot].data_ptr()
                slot_idx = self._slot_indices[step.slot].data_ptr()
                mean = step.nnz / max(1, step.row_hi - step.row_lo)
                args = self._args(rows=(step.row_lo, step.row_hi), k=k, out=V, ld_out=ldv,
                                  segs=[(slot_ptr, slot_idx, V + step.src_offset * ldv * w, ldv, step.row_lo)],
                                  accumulate=True, mean_degree=mean)
                items.append(("stream", step, args, events[l][i], (l, i)))
        # write-out
        if spec.emit_all_nodes:
            batch.append(self._args(rows=(0, st.num_nodes), k=k, out=aux, ld_out=max_k_aux,
                                    maps=[(D["inv_perm"].data_ptr(), V, ldv, 0, st.num_nodes)]))
        elif spec.direction == Direction.UP:
            rs = cv = None
            if spec.apply_endpoint_bias:
                if spec.init_mode == InitMode.XTX:
                    if "xtx_up" not in D:
                        raise RuntimeError("missing HIP up XTX endpoint bias")
                    rs = (D["xtx_up"].data_ptr(), 0)
                else:
                    if "vec_up" not in D:
                        raise RuntimeError("missing HIP up vector endpoint bias")
                    rs, cv = (D["vec_up"].data_ptr(), 0), io1
            wr = self._args(rows=(0, st.num_mutations), k=k, out=io0, ld_out=max_k,
                            segs=[(D["sel_ptr"].data_ptr(), D["sel_idx"].data_ptr(), V, ldv, 0)], row_scale=rs, col_vec=cv)
            self._tune(wr, (art.path, "write_sel", -1, k), private_out=True)
            batch.append(wr)
            if spec.use_miss:
                wr = self._args(rows=(0, st.num_mutations), k=k, out=io1, ld_out=max_k,
                                segs=[(D["miss_ptr"].data_ptr(), D["miss_idx"].data_ptr(), V, ldv, 0)])
                self._tune(wr, (art.path, "write_miss", -1, k), private_out=True)
                batch.append(wr)
        else:
            rs = cv = None
            suffix = "_ind" if spec.by_individual else ""
            if spec.apply_endpoint_bias:
                key = ("xtx_dn" if spec.init_mode == InitMode.XTX else "vec_dn") + suffix
                if key not in D:
                    raise RuntimeError(f"missing HIP down {spec.init_mode.value} endpoint bias")
                rs = (D[key].data_ptr(), 0)
                cv = io1 if spec.init_mode == InitMode.VECTOR else None
            if spec.by_individual:
                wr = self._args(rows=(0, st.num_individuals), k=k, out=io0, ld_out=max_k,
                                segs=[(D["ind_ptr"].data_ptr(), D["ind_idx"].data_ptr(), V, ldv, 0)],
                                row_scale=rs, col_vec=cv, mean_degree=float(st.ploidy))
                self._tune(wr, (art.path, "write_ind", -1, k), private_out=True)
                batch.append(wr)
            else:
                batch.append(self._args(rows=(0, st.num_samples), k=k, out=io0, ld_out=max_k,
                                        maps=[(D["out_map"].data_ptr(), V, ldv, 0, st.num_samples)], row_scale=rs, col_vec=cv))
        flush()
        return items

    # ---- execution -------------------------------------------------------

    def _copy_block_to_slot(self, step: _StreamStep, prev_compute_done, copy_done) -> None:
        """H2D copy of a streamed block into its ring slot on the slot's copy stream,
        after the previous compute that read the slot (``prev_compute_done``, None for
        the slot's first use in the sweep); ``copy_done`` is recorded after it. The
        counterpart of the reference's ``_copy_host_block_to_slot``
        (``triton/backend.py:958-968``)."""
        torch = _torch()
        cs = self._slot_copy_streams[step.slot]
        with torch.cuda.stream(cs):
            if prev_compute_done is not None:
                cs.wait_event(prev_compute_done)
            n_idx = int(step.host_indices.numel())
            n_ptr = int(step.host_indptr.numel())
            self._slot_indices[step.slot][:n_idx].copy_(step.host_indices, non_blocking=True)
            self._slot_indptr[step.slot][:n_ptr].copy_(step.host_indptr, non_blocking=True)
            copy_done.record(cs)

    def _launch_streamed_block(self, args: ffi.GrgHipGatherArgs, stream_ptr: int) -> None:
        """The accumulate launch of one streamed block (after its slot copy)."""
        self._lib.gather(args, stream_ptr)

    def _execute(self, art: _HipArtifact, spec: _CudaMatmulSpec, items) -> None:
        torch = _torch()
        with self._caller_root_scope():
            root = self._root_stream
            root_ptr = int(root.cuda_stream)
            has_stream = any(it[0] == "stream" for it in items)
            if has_stream and self._lib.stream_capturing(int(self._caller_stream.cuda_stream)):
                raise RuntimeError("HIP runtime: a sweep that streams blocks through the pinned ring cannot be captured "
                                   "into a hipGraph; plan a layout with every block resident to capture it")
            timer = self._sweep_timer
            if timer is not None:
                t_start, t_end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t_start.record(root)
            if has_stream:
                with torch.cuda.stream(root):
                    self._launch_event.record(root)
                for cs in self._slot_copy_streams:
                    cs.wait_event(self._launch_event)
            for it in items:
                if it[0] == "batch":
                    self._lib.gather_batch(it[1], it[2], root_ptr)
                    continue
                _, step, args, (copy_done, compute_done), _key = it
                prev = None
                if step.prev_in_slot is not None:
                    pl, pi = step.prev_in_slot
                    ev_list = art.up_events if spec.direction == Direction.UP else art.down_events
                    prev = ev_list[pl][pi][1]
                self._copy_block_to_slot(step, prev, copy_done)
                root.wait_event(copy_done)
                self._launch_streamed_block(args, root_ptr)
                with torch.cuda.stream(root):
                    compute_done.record(root)
            if timer is not None:
                t_end.record(root)
                timer.append((t_start, t_end))


__all__ = ["HipBatchedSweep", "HipRuntime", "launch_shape"]
