"""``plan_hip_layout``: device-memory plan for a set of ``.grg_spmv`` artifacts.

Same keyword surface, errors, warnings and residency rules as the reference's
``plan_triton_layout`` (``pygrgl_spmv/backends/triton/backend.py:343-561``):

* budget 0 means "exactly what full residency needs";
* ``allow_residency=False`` streams every block through ``ring_buffer_size``
  slots; otherwise everything is resident if it fits, else blocks are promoted
  to residency largest-first while ``fixed + resident + ring`` still fits;
* streamed blocks are assigned to slots round-robin in (artifact, owner, dst,
  src) order and each slot is sized to the largest block it serves;
* ``RuntimeWarning("requested ring_buffer_size=...")`` when the allocated slot
  count differs from the request; ``ValueError`` when nothing fits.

What differs is what is budgeted, because the HIP runtime allocates different
things: both sweep orientations are stored (no shared storage), selectors are
int32 CSR (forward for the up write-out, transposed for the down seed), the
resident blocks of each destination level are merged into one CSR whose row
pointers are counted under ``metadata``, and the per-level row-batch tables
and heavy-row lists get their own bounded category (``schedule``), as do the
tile tables of the LDS-staged level launches (``tiles``). ``bytes_total`` is an upper bound of what ``HipRuntime``
allocates for the layout.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
import warnings

import numpy as np

from pygrgl_spmv.backends import _hip_stream
from pygrgl_spmv.backends.base import BudgetItem, layout_struct_dtypes, sparse_structure_lengths, stored_block_shape
from pygrgl_spmv.backends.hip.plan import HipPlan, HipPlanPair
from pygrgl_spmv.backends.types import Direction
from pygrgl_spmv.grg import RuntimeRequirements
from pygrgl_spmv.grg.artifact import ArtifactScan, scan_grg_spmv

# Rows of gather degree above this (or above 8x their level's typical batch) get a
# workgroup of their own.
HEAVY_THRESHOLD = 64
# Rows of gather degree above SPLIT_MIN are cut into chunks of SPLIT_CHUNK entries,
# one workgroup each, whose FP64 partial rows the level launch then sums.
SPLIT_MIN = 8192
SPLIT_CHUNK = 2048


@dataclass
class HipBlockPlan:
    artifact_index: int
    owner: Direction
    shared: bool
    dst_level: int
    src_level: int
    store: object
    fmt: object
    stored_shape: tuple[int, int]
    nnz: int
    indices_dtype: np.dtype
    indptr_dtype: np.dtype
    indices_len: int
    indptr_len: int
    nbytes: int
    resident: bool = True
    slot: int | None = None


@dataclass(frozen=True)
class HipArtifactLayout:
    path: Path
    share_storage: bool
    up_owner: Direction | None
    down_owner: Direction | None
    blocks_up: tuple[HipBlockPlan, ...]
    blocks_down: tuple[HipBlockPlan, ...]


@dataclass(frozen=True)
class HipSlotPlan:
    indices_dtype: np.dtype
    indices_len: int
    indptr_dtype: np.dtype
    indptr_len: int

    @property
    def nbytes(self) -> int:
        return int(self.indices_len * self.indices_dtype.itemsize + self.indptr_len * self.indptr_dtype.itemsize)


@dataclass
class HipLayout:
    artifacts: tuple[HipArtifactLayout, ...]
    pair: HipPlanPair
    dtype: np.dtype
    requirements: RuntimeRequirements
    device: int
    stream_ptr: int
    stream_owner: object | None
    allow_residency: bool
    requested_ring_buffer_size: int
    allocated_ring_buffer_size: int
    vram_budget_bytes: int
    max_num_samples: int
    max_num_mutations: int
    max_num_nodes: int
    max_num_individuals: int
    max_selector_nnz: int
    max_levels: int
    max_rows_by_level: tuple[int, ...]
    max_up_ops_by_level: tuple[int, ...]
    max_down_ops_by_level: tuple[int, ...]
    slot_plans: tuple[HipSlotPlan, ...]
    budget_items: tuple[BudgetItem, ...]
    required_budget_for_full_residency: int
    bytes_by_category: dict = field(default_factory=dict)
    bytes_total: int = 0
    heavy_threshold: int = HEAVY_THRESHOLD
    schedule_capacity: tuple[int, ...] = ()
    tiles: bool = False
    # True when the caller capped device memory; an uncapped layout stores its full-residency size above.
    budgeted: bool = False


def _resolve_artifacts(artifacts) -> tuple[Path, ...]:
    paths = tuple(Path(p) for p in artifacts)
    if not paths:
        raise ValueError("artifacts must be non-empty")
    for p in paths:
        if p.suffix != ".grg_spmv":
            raise ValueError(f"planner expects .grg_spmv artifacts, got {p}")
        if not p.exists():
            raise FileNotFoundError(p)
    return paths


def _block_plan(artifact_index: int, owner: Direction, block, plan: HipPlan) -> HipBlockPlan:
    nrows, ncols = stored_block_shape(block.shape[0], block.shape[1], store=plan.store)
    ptr_dt, idx_dt = layout_struct_dtypes(plan.fmt, nrows=nrows, ncols=ncols, nnz=block.nnz)
    ptr_len, idx_len = sparse_structure_lengths(plan.fmt, nrows=nrows, ncols=ncols, nnz=block.nnz)
    return HipBlockPlan(
        artifact_index=artifact_index, owner=owner, shared=False,
        dst_level=int(block.dst_level), src_level=int(block.src_level),
        store=plan.store, fmt=plan.fmt, stored_shape=(nrows, ncols), nnz=int(block.nnz),
        indices_dtype=np.dtype(idx_dt), indptr_dtype=np.dtype(ptr_dt),
        indices_len=int(idx_len), indptr_len=int(ptr_len),
        nbytes=int(ptr_len * np.dtype(ptr_dt).itemsize + idx_len * np.dtype(idx_dt).itemsize),
    )


def _assign_slots(blocks: list[HipBlockPlan], ring: int) -> tuple[tuple[HipSlotPlan, ...], int]:
    for b in blocks:
        b.slot = None
    streamed = sorted((b for b in blocks if not b.resident), key=lambda b: (b.artifact_index, b.owner.value, b.dst_level, b.src_level))
    if not streamed:
        return (), 0
    n_slots = min(int(ring), len(streamed))
    if n_slots < 1:
        raise ValueError("ring_buffer_size must be >= 1 when any HIP block is streamed")
    idx_dt = [np.dtype(np.int32)] * n_slots
    ptr_dt = [np.dtype(np.int32)] * n_slots
    idx_len = [0] * n_slots
    ptr_len = [0] * n_slots
    for i, b in enumerate(streamed):
        s = i % n_slots
        b.slot = s
        if b.indices_dtype == np.dtype(np.int64):
            idx_dt[s] = np.dtype(np.int64)
        if b.indptr_dtype == np.dtype(np.int64):
            ptr_dt[s] = np.dtype(np.int64)
        idx_len[s] = max(idx_len[s], b.indices_len)
        ptr_len[s] = max(ptr_len[s], b.indptr_len)
    slots = tuple(HipSlotPlan(idx_dt[s], idx_len[s], ptr_dt[s], ptr_len[s]) for s in range(n_slots))
    return slots, int(sum(s.nbytes for s in slots))


def _promotion_key(b: HipBlockPlan):
    return (-b.nbytes, -b.artifact_index, -b.dst_level, -b.src_level, 0 if b.owner == Direction.UP else 1)


def _warn_ring(requested: int, allocated: int) -> None:
    if allocated == requested:
        return
    if allocated == 0 and requested > 0:
        msg = f"requested ring_buffer_size={requested}, but all sparse blocks fit resident so no ring slots were allocated"
    else:
        msg = f"requested ring_buffer_size={requested}, but only {allocated} ring slot(s) were allocated"
    warnings.warn(msg, RuntimeWarning, stacklevel=3)


def enabled_max_k(pair: HipPlanPair, req: RuntimeRequirements) -> int:
    ks = ([int(req.max_k_up)] if pair.plan_up is not None else []) + ([int(req.max_k_down)] if pair.plan_down is not None else [])
    if not ks:
        raise ValueError("HIP layout requires at least one configured direction plan")
    return max(ks)


def side_io_rows(pair: HipPlanPair, req: RuntimeRequirements, *, max_num_mutations: int, max_num_nodes: int) -> int:
    need_miss = (pair.plan_down is not None and req.need_down_miss_input) or (pair.plan_up is not None and req.need_up_miss_output)
    return max(
        int(max_num_mutations) if need_miss else 0,
        1 if req.need_init_vector else 0,
        int(max_num_nodes) if req.need_init_matrix else 0,
    )


def _selector_bytes(scan: ArtifactScan, pair: HipPlanPair, req: RuntimeRequirements) -> int:
    M, N = scan.num_mutations, scan.num_nodes
    total = 0
    if pair.plan_up is not None:
        total += (M + 1 + scan.selector_mut_nnz)
        if req.need_up_miss_output:
            total += (M + 1 + scan.selector_miss_nnz)
    if pair.plan_down is not None:
        total += (N + 1 + scan.selector_mut_nnz)
        if req.need_down_miss_input:
            total += (N + 1 + scan.selector_miss_nnz)
    return 4 * total


def _metadata_bytes(scan: ArtifactScan, pair: HipPlanPair, req: RuntimeRequirements, dtype: np.dtype) -> int:
    S, M, N, I = scan.num_samples, scan.num_mutations, scan.num_nodes, scan.num_individuals
    w = dtype.itemsize
    total = 4 * (2 * N + 3 * S + I + 1 + S)  # node maps (2N), sample input/output maps (3S), individual -> samples CSR
    total += 4 * (N + 1) * ((pair.plan_up is not None) + (pair.plan_down is not None))  # merged level row pointers
    if req.need_init_vector:
        total += w * ((M if pair.plan_up is not None else 0) + ((S + I) if pair.plan_down is not None else 0))
    if req.need_init_xtx and scan.has_xtx_bias:
        total += w * ((M if pair.plan_up is not None else 0) + ((S + I) if pair.plan_down is not None else 0))
    return int(total)


def _gather_entries_bound(scan: ArtifactScan) -> int:
    """Upper bound of the gather entries of any one sweep configuration."""
    return int(scan.num_edges + scan.selector_mut_nnz + scan.selector_miss_nnz)


def split_items_bound(scan: ArtifactScan) -> int:
    """Upper bound of the split items of one level launch: a split row of degree
    d > SPLIT_MIN has ceil(d / SPLIT_CHUNK) < d / SPLIT_CHUNK + d / SPLIT_MIN items."""
    e = _gather_entries_bound(scan)
    return int(e // SPLIT_CHUNK + e // SPLIT_MIN + 1)


def schedule_capacity_for(scan: ArtifactScan, pair: HipPlanPair, req: RuntimeRequirements) -> int:
    """int32 entries reserved for heavy-row lists and split-row tables.

    Per sweep configuration (up; down; down with missing-data input): at most
    ``num_nodes`` heavy rows, one item offset per heavy row plus one per level,
    and the split items of all levels.
    """
    configs = (pair.plan_up is not None) + (pair.plan_down is not None) * (2 if req.need_down_miss_input else 1)
    per_config = 2 * scan.num_nodes + scan.num_levels + split_items_bound(scan)
    return int(configs * per_config)


PANEL_ROWS = 1024         # rows of one panel of the panel launches
PANEL_STAGED = 255        # source rows staged per panel pass (8-bit positions)
PANEL_PASS_ENTRIES = 8000  # entries of one panel pass (its positions sit in 8 KiB of LDS)
GROUP_ROWS = 16       # rows of one group of the row-group launches (FP64 values)


def tile_bytes_for(scan: ArtifactScan, pair: HipPlanPair, req: RuntimeRequirements, dtype=np.float64) -> int:
    """Upper bound of the tile tables of the tiled level launches and (FP64 values) of
    the row groups of the row-group launches. (The panels of the panel launches, off by
    default, are not reserved here: a runtime builds them only under a layout without
    a VRAM budget.)

    Per sweep configuration (up; down; down with missing-data input): a uint16
    staged index per gather entry (plus 16 bytes of padding), at most one uint32
    staged source per entry, a uint32 merged entry key per entry (configurations
    of more than one segment),
    int32 merged row pointers, and int32 tile row and staged-source offsets (at
    most one tile per row plus one per level).
    """
    N, L, E = scan.num_nodes, scan.num_levels, scan.num_edges
    entries = []  # (entries, merged keys kept: more than one segment)
    if pair.plan_up is not None:
        entries.append((E, False))
    if pair.plan_down is not None:
        entries.append((E + scan.selector_mut_nnz, True))
        if req.need_down_miss_input:
            entries.append((E + scan.selector_mut_nnz + scan.selector_miss_nnz, True))
    tiles = sum(6 * e + (4 * e if keyed else 0) + 16 + 4 * (N + 1) + 8 * (N + L + 1) for e, keyed in entries)
    # row groups (no more than one per GROUP_ROWS rows plus one per level): at most one
    # uint32 key and uint16 mask per entry plus three padding entries per group, an
    # int64 offset and a uint16 store mask per group
    grp = 0
    if np.dtype(dtype) == np.float64:
        groups = N // GROUP_ROWS + L + 1
        for e, _ in entries:
            grp += 6 * (e + 3 * groups) + 10 * (groups + 1) + 32
    return int(tiles + grp)


def plan_hip_layout(
    *,
    artifacts,
    pair: HipPlanPair,
    dtype,
    requirements: RuntimeRequirements,
    vram_budget_bytes: int,
    ring_buffer_size: int,
    allow_residency: bool = True,
    device,
    stream,
    tiles: bool = True,
) -> HipLayout:
    """Plan device memory for ``artifacts`` (no device allocation happens here).

    ``tiles`` (no reference counterpart) reserves the tile tables of the
    LDS-staged level launches (category ``tiles``, an upper bound); without it
    the runtime runs every level with the per-row gather path."""
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError(f"HIP runtime supports only float32/float64, got {dtype}")
    ring = int(ring_buffer_size)
    budget = int(vram_budget_bytes)
    if ring < 0:
        raise ValueError("ring_buffer_size must be >= 0")
    if budget < 0:
        raise ValueError("vram_budget_bytes must be >= 0")
    paths = _resolve_artifacts(artifacts)
    device_id = _hip_stream.parse_hip_device(device)
    stream_ptr, stream_owner = _hip_stream.parse_hip_stream(stream)
    if stream_ptr != 0:
        sdev = _hip_stream.hip_stream_device(stream_ptr)
        if sdev != device_id:
            raise ValueError(f"HIP stream device {sdev} does not match requested HIP device {device_id}")
    allow_residency = bool(allow_residency)
    scans = tuple(scan_grg_spmv(p) for p in paths)
    max_levels = max(s.num_levels for s in scans)
    max_rows_by_level = tuple(max((s.level_sizes[l] if l < s.num_levels else 0) for s in scans) for l in range(max_levels))
    max_up_ops = tuple(max(sum(1 for b in s.blocks if b.nnz > 0 and b.dst_level == l) for s in scans) for l in range(max_levels))
    max_down_ops = tuple(max(sum(1 for b in s.blocks if b.nnz > 0 and b.src_level == l) for s in scans) for l in range(max_levels))
    S_max = max(s.num_samples for s in scans)
    M_max = max(s.num_mutations for s in scans)
    N_max = max(s.num_nodes for s in scans)
    I_max = max(s.num_individuals for s in scans)
    sel_max = max(max(s.selector_mut_nnz, s.selector_miss_nnz) for s in scans)

    selector_bytes = metadata_bytes = xtx_bytes = heavy_bytes = tile_bytes = 0
    heavy_caps = []
    items_max = 0
    for s in scans:
        selector_bytes += _selector_bytes(s, pair, requirements)
        metadata_bytes += _metadata_bytes(s, pair, requirements, dtype)
        if requirements.need_init_xtx and s.has_individual_coals:
            xtx_bytes += s.num_nodes * dtype.itemsize
        if tiles:
            tile_bytes += tile_bytes_for(s, pair, requirements, dtype)
        cap = schedule_capacity_for(s, pair, requirements)
        heavy_caps.append(cap)
        heavy_bytes += 4 * cap
        items_max = max(items_max, split_items_bound(s))
    max_k = enabled_max_k(pair, requirements)
    scratch_bytes = 8 * items_max * max_k  # FP64 partial rows of split rows, reused level by level
    state_bytes = N_max * max_k * dtype.itemsize
    io0_rows = max(S_max, M_max)
    io1_rows = side_io_rows(pair, requirements, max_num_mutations=M_max, max_num_nodes=N_max)
    aux_rows = max(N_max, S_max, M_max, sel_max)
    dense_bytes = (io0_rows + io1_rows + aux_rows) * max_k * dtype.itemsize

    blocks: list[HipBlockPlan] = []
    planned = []
    for ai, (path, scan) in enumerate(zip(paths, scans)):
        ups = () if pair.plan_up is None else tuple(_block_plan(ai, Direction.UP, b, pair.plan_up) for b in scan.blocks if b.nnz > 0)
        downs = () if pair.plan_down is None else tuple(_block_plan(ai, Direction.DOWN, b, pair.plan_down) for b in scan.blocks if b.nnz > 0)
        blocks.extend(ups)
        blocks.extend(downs)
        planned.append(HipArtifactLayout(
            path=path, share_storage=False,
            up_owner=Direction.UP if pair.plan_up is not None else None,
            down_owner=Direction.DOWN if pair.plan_down is not None else None,
            blocks_up=ups, blocks_down=downs,
        ))

    fixed = int(selector_bytes + metadata_bytes + xtx_bytes + heavy_bytes + tile_bytes + dense_bytes + state_bytes + scratch_bytes)
    full_resident = int(sum(b.nbytes for b in blocks))
    required_full = fixed + full_resident
    effective = required_full if budget == 0 else budget
    if not allow_residency:
        if blocks and ring < 1:
            raise ValueError("ring_buffer_size must be >= 1 when any HIP block is streamed")
        for b in blocks:
            b.resident = False
        slots, ring_bytes = _assign_slots(blocks, ring) if blocks else ((), 0)
        resident_bytes = 0
        total = fixed + ring_bytes
        if total > effective:
            raise ValueError(f"HIP layout requires at least {total} owned device bytes, exceeds vram_budget_bytes={effective}")
        if len(slots) != ring:
            _warn_ring(ring, len(slots))
    elif required_full <= effective:
        for b in blocks:
            b.resident = True
            b.slot = None
        slots, ring_bytes, resident_bytes, total = (), 0, full_resident, required_full
        if ring > 0:
            _warn_ring(ring, 0)
    else:
        if ring < 1:
            raise ValueError("ring_buffer_size must be >= 1 when any HIP block is streamed")
        for b in blocks:
            b.resident = False
        slots, ring_bytes = _assign_slots(blocks, ring)
        if fixed + ring_bytes > effective:
            raise ValueError(f"HIP layout requires at least {fixed + ring_bytes} owned device bytes, exceeds vram_budget_bytes={effective}")
        for b in sorted(blocks, key=_promotion_key):
            b.resident = True
            resident_bytes = sum(x.nbytes for x in blocks if x.resident)
            slots, ring_bytes = _assign_slots(blocks, ring)
            if fixed + resident_bytes + ring_bytes > effective:
                b.resident = False
        resident_bytes = int(sum(x.nbytes for x in blocks if x.resident))
        slots, ring_bytes = _assign_slots(blocks, ring)
        total = fixed + resident_bytes + ring_bytes
        if len(slots) != ring:
            _warn_ring(ring, len(slots))
    if total > effective:
        raise ValueError(f"HIP layout requires {total} owned device bytes, exceeds vram_budget_bytes={effective}")

    by_cat = {
        "resident_sparse": int(resident_bytes),
        "ring_slots": int(ring_bytes),
        "selectors": int(selector_bytes),
        "metadata": int(metadata_bytes),
        "xtx_bias": int(xtx_bytes),
        "schedule": int(heavy_bytes),
        "tiles": int(tile_bytes),
        "dense_arena": int(dense_bytes),
        "state": int(state_bytes),
        "scratch": int(scratch_bytes),
    }
    items = [BudgetItem(kind="fixed", name=name, nbytes=int(by_cat[name]))
             for name in ("selectors", "metadata", "xtx_bias", "schedule", "tiles", "dense_arena", "state", "scratch")]
    for ai, al in enumerate(planned):
        for b in (*al.blocks_up, *al.blocks_down):
            if b.resident:
                items.append(BudgetItem(kind="resident_sparse", name="resident_sparse", nbytes=int(b.nbytes),
                                        artifact_index=ai, dst_level=b.dst_level, src_level=b.src_level))
    for si, sl in enumerate(slots):
        items.append(BudgetItem(kind="ring_slot", name="ring_slots", nbytes=int(sl.nbytes), slot=si))
    return HipLayout(
        artifacts=tuple(planned), pair=pair, dtype=dtype, requirements=requirements,
        device=device_id, stream_ptr=int(stream_ptr), stream_owner=stream_owner,
        allow_residency=allow_residency, requested_ring_buffer_size=ring, allocated_ring_buffer_size=len(slots),
        vram_budget_bytes=int(effective), max_num_samples=S_max, max_num_mutations=M_max, max_num_nodes=N_max,
        max_num_individuals=I_max, max_selector_nnz=sel_max, max_levels=max_levels, max_rows_by_level=max_rows_by_level,
        max_up_ops_by_level=max_up_ops, max_down_ops_by_level=max_down_ops, slot_plans=slots,
        budget_items=tuple(i for i in items if i.nbytes > 0), required_budget_for_full_residency=int(required_full),
        bytes_by_category=by_cat, bytes_total=int(sum(i.nbytes for i in items)),
        heavy_threshold=HEAVY_THRESHOLD, schedule_capacity=tuple(heavy_caps), tiles=bool(tiles),
        budgeted=budget > 0,
    )


__all__ = ["HEAVY_THRESHOLD", "PANEL_ROWS", "PANEL_STAGED", "SPLIT_CHUNK", "SPLIT_MIN", "split_items_bound", "HipArtifactLayout", "HipBlockPlan", "HipLayout", "HipSlotPlan", "enabled_max_k", "plan_hip_layout", "side_io_rows", "tile_bytes_for"]
