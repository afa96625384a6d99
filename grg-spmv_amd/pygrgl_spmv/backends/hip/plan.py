"""HIP plan types (counterpart of the reference's ``triton/plan.py:41-94``).

The HIP sweeps are pull-only, so each direction stores the adjacency its
destination rows read:

* up (samples -> mutations) reads each parent's children: ``A`` in CSR,
  i.e. ``store="N", fmt="CSR"``;
* down (mutations -> samples) reads each node's parents: ``A.T`` in CSR,
  i.e. ``store="T", fmt="CSR"`` (the CSC of ``A``).

The two orientations never share storage, and the planner budgets both. The
reference's default down plan (``store="T", fmt="CSC"``, shared with up) implies
a scatter with atomics, which this backend does not run; it is rejected with
a message naming the supported plan.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Mapping

from pygrgl_spmv.backends.types import (
    SparseFormat,
    StoredMatrix,
    parse_sparse_format,
    parse_store,
    transpose_compatible_format,
)


@dataclass(frozen=True)
class HipPlan:
    store: StoredMatrix
    fmt: SparseFormat

    def __post_init__(self) -> None:
        object.__setattr__(self, "store", parse_store(self.store))
        object.__setattr__(self, "fmt", parse_sparse_format(self.fmt))
        if self.fmt != SparseFormat.CSR:
            raise ValueError(f"HIP backend stores CSR blocks (pull sweeps); got fmt={self.fmt.value}")

    @classmethod
    def from_dict(cls, value: Mapping[str, object]) -> "HipPlan":
        allowed = {"store", "fmt", "scratch"}
        extra = sorted(set(value) - allowed)
        if extra:
            raise ValueError(f"unknown HipPlan field(s): {extra}")
        scratch = str(value.get("scratch", "none")).strip().lower()
        if scratch not in ("", "none"):
            raise ValueError("HIP plans take no scratch levels: every level is one fused launch that writes each row once")
        return cls(store=parse_store(value["store"]), fmt=parse_sparse_format(value["fmt"]))

    def can_share_storage_with(self, other: "HipPlan") -> bool:
        if self.store == other.store:
            return self.fmt == other.fmt
        return transpose_compatible_format(self.fmt) == other.fmt


@dataclass(frozen=True)
class HipPlanPair:
    plan_up: HipPlan | None
    plan_down: HipPlan | None

    def __post_init__(self) -> None:
        if self.plan_up is None and self.plan_down is None:
            raise ValueError("at least one of plan_up/plan_down must be provided")
        if self.plan_up is not None and self.plan_up.store != StoredMatrix.N:
            raise ValueError(f"HIP plan_up expects store=N (parent-major CSR), got {self.plan_up.store.value}")
        if self.plan_down is not None and self.plan_down.store != StoredMatrix.T:
            raise ValueError(f"HIP plan_down expects store=T (child-major CSR), got {self.plan_down.store.value}")

    @classmethod
    def from_dicts(cls, plan_up: Mapping[str, object] | None, plan_down: Mapping[str, object] | None) -> "HipPlanPair":
        return cls(
            plan_up=None if plan_up is None else HipPlan.from_dict(plan_up),
            plan_down=None if plan_down is None else HipPlan.from_dict(plan_down),
        )

    @classmethod
    def default(cls, *, up: bool = True, down: bool = True) -> "HipPlanPair":
        return cls(
            plan_up=HipPlan(store="N", fmt="CSR") if up else None,
            plan_down=HipPlan(store="T", fmt="CSR") if down else None,
        )


__all__ = ["HipPlan", "HipPlanPair"]
