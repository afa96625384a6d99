"""``plan_hip_layout`` contract (CPU, fake device).

Restates the planner tests of the reference's ``runtime/test_triton.py:230-505``
and ``_streaming_cases.py:384-448`` (three-block transition table),
``runtime/test_directional_layouts.py:189-249`` (one-sided layouts, disabled
direction requirements, dense arena of the sparse-selector toy) and
``runtime/test_plans.py`` (plan parsing). HIP stores both sweep orientations,
so a full up+down pair owns six blocks of a three-block artifact; the
three-block table is checked on one-sided pairs (three owner blocks each) and
the six-block thresholds on the full pair.
"""

from __future__ import annotations

from dataclasses import dataclass
import warnings

import numpy as np
import pytest
import scipy.sparse as sp

from pygrgl_spmv.backends.hip import HipPlan, HipPlanPair
from pygrgl_spmv.grg.artifact import save_grg_spmv
from pygrgl_spmv.grg.sparse import binary_csr_from_parts
from tests._builders import build_hip_layout, full_requirements, synthetic_state, write_band_artifact

pytestmark = pytest.mark.usefixtures("fake_hip_device")

KEYS = ((1, 0), (2, 0), (2, 1))


@dataclass(frozen=True)
class Mode:
    name: str
    ring: int
    budget_blocks: int
    resident: tuple
    streamed: tuple
    slots: int
    warn: bool


MODES = (
    Mode("ring0-resident", 0, 3, KEYS, (), 0, False),
    Mode("ring1-stream", 1, 1, (), KEYS, 1, False),
    Mode("ring1-hybrid", 1, 2, ((2, 1),), ((1, 0), (2, 0)), 1, False),
    Mode("ring1-resident", 1, 3, KEYS, (), 0, True),
    Mode("ring2-stream", 2, 2, (), KEYS, 2, False),
    Mode("ring2-resident", 2, 3, KEYS, (), 0, True),
    Mode("ring3-resident", 3, 3, KEYS, (), 0, True),
)
BY_NAME = {m.name: m for m in MODES}
BELOW = {"ring0-resident": None, "ring1-stream": None, "ring1-hybrid": "ring1-stream", "ring1-resident": "ring1-hybrid",
         "ring2-stream": None, "ring2-resident": "ring2-stream", "ring3-resident": None}


@pytest.fixture(scope="module")
def band(tmp_path_factory):
    return write_band_artifact(tmp_path_factory.mktemp("band"), "band-n64", n=64, bandwidth=8)


def _pair(side: str) -> HipPlanPair:
    return HipPlanPair.default(up=side in ("up", "both"), down=side in ("down", "both"))


def _layout(art, *, side="up", k=2, ring=0, budget=10**12, allow_residency=True):
    return build_hip_layout([art], pair=_pair(side), requirements=full_requirements(max_k_up=k, max_k_down=k),
                            ring_buffer_size=ring, vram_budget_bytes=budget, allow_residency=allow_residency)


def _components(art, side, k):
    lay = _layout(art, side=side, k=k)
    sizes = {b.nbytes for a in lay.artifacts for b in (*a.blocks_up, *a.blocks_down)}
    assert len(sizes) == 1
    fixed = lay.bytes_total - lay.bytes_by_category["resident_sparse"]
    count = sum(1 for a in lay.artifacts for _ in (*a.blocks_up, *a.blocks_down))
    return fixed, sizes.pop(), count


def _keys(lay, resident: bool):
    return tuple(sorted((b.dst_level, b.src_level) for a in lay.artifacts for b in (*a.blocks_up, *a.blocks_down) if b.resident == resident))


def _build_mode(art, side, k, mode, budget):
    if mode.warn:
        with pytest.warns(RuntimeWarning, match="requested ring_buffer_size"):
            return _layout(art, side=side, k=k, ring=mode.ring, budget=budget)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        return _layout(art, side=side, k=k, ring=mode.ring, budget=budget)


def _check_mode(lay, mode, budget, fixed, block):
    assert lay.requested_ring_buffer_size == mode.ring
    assert lay.allocated_ring_buffer_size == mode.slots
    assert _keys(lay, True) == tuple(sorted(mode.resident))
    assert _keys(lay, False) == tuple(sorted(mode.streamed))
    assert sum(1 for s in lay.slot_plans if s.nbytes > 0) == mode.slots
    assert len([i for i in lay.budget_items if i.kind == "resident_sparse"]) == len(mode.resident)
    assert len([i for i in lay.budget_items if i.kind == "ring_slot"]) == mode.slots
    assert sum(i.nbytes for i in lay.budget_items) == lay.bytes_total
    assert lay.required_budget_for_full_residency == fixed + 3 * block
    assert lay.bytes_total <= budget <= lay.required_budget_for_full_residency


@pytest.mark.parametrize("side", ["up", "down"])
@pytest.mark.parametrize("k", [1, 2])
@pytest.mark.parametrize("mode", MODES, ids=[m.name for m in MODES])
def test_three_block_transitions(band, side, k, mode):
    fixed, block, count = _components(band, side, k)
    assert count == 3
    budget = fixed + mode.budget_blocks * block
    lay = _build_mode(band, side, k, mode, budget)
    _check_mode(lay, mode, budget, fixed, block)
    below = BELOW[mode.name]
    if below is None:
        with pytest.raises(ValueError):
            _layout(band, side=side, k=k, ring=mode.ring, budget=budget - 1)
        return
    lower = _build_mode(band, side, k, BY_NAME[below], budget - 1)
    _check_mode(lower, BY_NAME[below], budget - 1, fixed, block)


def test_full_pair_owns_both_orientations(band):
    fixed, block, count = _components(band, "both", 2)
    assert count == 6
    lay = _layout(band, side="both", ring=2, budget=fixed + 4 * block)
    assert lay.allocated_ring_buffer_size == 2
    assert len(_keys(lay, True)) == 2 and len(_keys(lay, False)) == 4
    owners = {(b.owner.value, b.slot) for a in lay.artifacts for b in (*a.blocks_up, *a.blocks_down) if not b.resident}
    assert {s for _, s in owners} == {0, 1}


@pytest.mark.parametrize("budget,ring,case", [(0, 0, "full"), (0, 1, "full_warn"), ("tight", 0, "ring_error"), ("tight", 1, "streamed")])
def test_zero_budget_sentinel_and_tight_ring(band, budget, ring, case):
    full = _layout(band, side="both", budget=0)
    block = min(b.nbytes for a in full.artifacts for b in (*a.blocks_up, *a.blocks_down))
    b = full.required_budget_for_full_residency - block if budget == "tight" else budget
    if case == "ring_error":
        with pytest.raises(ValueError, match="ring_buffer_size"):
            _layout(band, side="both", ring=ring, budget=b)
        return
    if case == "full_warn":
        with pytest.warns(RuntimeWarning, match="all sparse blocks fit resident"):
            lay = _layout(band, side="both", ring=ring, budget=b)
    else:
        lay = _layout(band, side="both", ring=ring, budget=b)
    if case.startswith("full"):
        assert _keys(lay, False) == ()
        assert lay.allocated_ring_buffer_size == 0
        assert lay.vram_budget_bytes == lay.required_budget_for_full_residency == lay.bytes_total
        assert not lay.budgeted
    else:
        assert lay.budgeted
        assert lay.allocated_ring_buffer_size == 1
        assert lay.bytes_total <= lay.vram_budget_bytes
        assert lay.bytes_total < lay.required_budget_for_full_residency


def test_negative_budget_and_ring_rejected(band):
    with pytest.raises(ValueError, match="vram_budget_bytes"):
        _layout(band, budget=-1)
    with pytest.raises(ValueError, match="ring_buffer_size"):
        _layout(band, ring=-1)


@pytest.mark.parametrize("ring", [1, 2, 3])
def test_allow_residency_false_streams_everything(band, ring):
    fixed, block, _ = _components(band, "up", 2)
    lay = _layout(band, side="up", ring=ring, budget=fixed + 3 * block, allow_residency=False)
    assert lay.allow_residency is False
    assert lay.allocated_ring_buffer_size == ring
    assert _keys(lay, True) == () and _keys(lay, False) == KEYS
    assert lay.bytes_by_category["resident_sparse"] == 0
    assert lay.bytes_total == fixed + ring * block
    assert lay.required_budget_for_full_residency == fixed + 3 * block


def test_allow_residency_false_rejects_ring_zero(band):
    with pytest.raises(ValueError, match="ring_buffer_size"):
        _layout(band, ring=0, allow_residency=False)


def test_allow_residency_false_warns_when_ring_exceeds_blocks(band):
    fixed, block, _ = _components(band, "up", 2)
    with pytest.warns(RuntimeWarning, match="requested ring_buffer_size"):
        lay = _layout(band, side="up", ring=4, budget=fixed + 3 * block, allow_residency=False)
    assert lay.allocated_ring_buffer_size == 3


def test_one_sided_layout_drops_unused_storage(band):
    full = _layout(band, side="both", k=4)
    up_only = _layout(band, side="up", k=4)
    assert up_only.bytes_total < full.bytes_total
    assert up_only.bytes_by_category["resident_sparse"] < full.bytes_by_category["resident_sparse"]


@pytest.mark.parametrize("side", ["up", "down"])
def test_one_sided_layout_ignores_disabled_direction_requirements(band, side):
    def req(disabled_k, disabled_miss):
        if side == "up":
            return full_requirements(max_k_up=2, max_k_down=disabled_k, need_down_miss_input=disabled_miss, need_up_miss_output=False)
        return full_requirements(max_k_up=disabled_k, max_k_down=2, need_down_miss_input=False, need_up_miss_output=disabled_miss)

    benign = build_hip_layout([band], pair=_pair(side), requirements=req(1, False))
    hostile = build_hip_layout([band], pair=_pair(side), requirements=req(1024, True), vram_budget_bytes=benign.bytes_total)
    assert hostile.bytes_total == benign.bytes_total
    assert hostile.required_budget_for_full_residency == benign.required_budget_for_full_residency
    assert hostile.bytes_by_category == benign.bytes_by_category


def test_dense_arena_includes_mutation_rows_with_sparse_selector(tmp_path):
    """Reference test_directional_layouts.py:232-249: (io0 5 + io1 1 + aux 5) rows x k=2 x 8 bytes."""
    sdt = np.dtype(np.int32)
    sel_mut = binary_csr_from_parts(indices=np.asarray([1], dtype=sdt), indptr=np.asarray([0, 1, 1, 1, 1, 1], dtype=sdt), shape=(5, 2), shared_data=True)
    sel_miss = binary_csr_from_parts(indices=np.empty(0, dtype=sdt), indptr=np.zeros(6, dtype=sdt), shape=(5, 2), shared_data=True)
    state = synthetic_state(blocks=[[], [sp.csr_matrix((1, 1), dtype=np.bool_)]], level_offsets=np.asarray([0, 1, 2], dtype=sdt),
                            num_samples=1, num_mutations=5, num_nodes=2, sel_mut=sel_mut, sel_miss=sel_miss)
    path = tmp_path / "sparse-mutation-selector.grg_spmv"
    save_grg_spmv(state, path)
    req = full_requirements(max_k_up=2, max_k_down=2, need_down_miss_input=False, need_up_miss_output=False,
                            need_init_vector=True, need_init_matrix=False, need_init_xtx=False)
    lay = build_hip_layout([path], requirements=req)
    assert lay.bytes_by_category["dense_arena"] == (5 + 1 + 5) * 2 * 8


def test_layout_rejects_bad_inputs(band, tmp_path):
    with pytest.raises(ValueError, match="float32/float64"):
        build_hip_layout([band], dtype=np.int32)
    with pytest.raises(ValueError, match="non-empty"):
        build_hip_layout([])
    with pytest.raises(ValueError, match=".grg_spmv"):
        build_hip_layout([tmp_path / "x.npz"])
    with pytest.raises(FileNotFoundError):
        build_hip_layout([tmp_path / "missing.grg_spmv"])
    with pytest.raises(ValueError, match="out of range"):
        build_hip_layout([band], device=3)
    with pytest.raises(TypeError):
        build_hip_layout([band], device=True)
    with pytest.raises(TypeError):
        build_hip_layout([band], stream=object())


def test_plan_types():
    assert HipPlan.from_dict({"store": "N", "fmt": "CSR"}) == HipPlan(store="N", fmt="CSR")
    with pytest.raises(ValueError, match="unknown HipPlan"):
        HipPlan.from_dict({"store": "N", "fmt": "CSR", "algo": 1})
    with pytest.raises(ValueError, match="CSR"):
        HipPlan(store="T", fmt="CSC")
    with pytest.raises(ValueError, match="scratch"):
        HipPlan.from_dict({"store": "N", "fmt": "CSR", "scratch": "1"})
    with pytest.raises(ValueError, match="at least one"):
        HipPlanPair(plan_up=None, plan_down=None)
    with pytest.raises(ValueError, match="store=N"):
        HipPlanPair(plan_up=HipPlan(store="T", fmt="CSR"), plan_down=None)
    with pytest.raises(ValueError, match="store=T"):
        HipPlanPair(plan_up=None, plan_down=HipPlan(store="N", fmt="CSR"))
    pair = HipPlanPair.from_dicts({"store": "N", "fmt": "CSR"}, {"store": "T", "fmt": "CSR"})
    assert not pair.plan_up.can_share_storage_with(pair.plan_down)
    assert HipPlan(store="N", fmt="CSR").can_share_storage_with(HipPlan(store="N", fmt="CSR"))
