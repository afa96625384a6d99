"""Mixed-model solve on the HIP GRG sweeps: the standardized kinship operator, batched
FP64 conjugate gradients, variance components and leave-one-GRG-out solves, top
principal components and the learned CG warm start."""

from pygrgl_spmv.lmm.cg import CgResult, conj_grad, shifted
from pygrgl_spmv.lmm.operator import StandardizedGrgOperator, center
from pygrgl_spmv.lmm.pcs import Eigenpairs, top_eigenpairs
from pygrgl_spmv.lmm.simulate import Phenotypes, simulate
from pygrgl_spmv.lmm.variance import McScaling, VarianceFit, fit_variance_components, mc_scaling, solve_loco
from pygrgl_spmv.lmm.warmstart import Basis, WarmStartNet, closed_form_coefficients, features, start_from_coefficients, targets

__all__ = [
    "Basis", "CgResult", "Eigenpairs", "McScaling", "Phenotypes", "VarianceFit", "fit_variance_components", "mc_scaling",
    "solve_loco", "StandardizedGrgOperator", "WarmStartNet", "center",
    "closed_form_coefficients", "conj_grad", "features", "shifted", "simulate", "start_from_coefficients", "targets",
    "top_eigenpairs",
]
