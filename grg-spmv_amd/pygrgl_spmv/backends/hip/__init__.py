"""MI355X HIP backend: ``plan_hip_layout`` / ``HipRuntime`` (drop-in for the Triton and cuSPARSE backends)."""

from pygrgl_spmv.backends.hip.layout import HipLayout, plan_hip_layout
from pygrgl_spmv.backends.hip.plan import HipPlan, HipPlanPair
from pygrgl_spmv.backends.hip.runtime import HipRuntime

__all__ = ["HipLayout", "HipPlan", "HipPlanPair", "HipRuntime", "plan_hip_layout"]
