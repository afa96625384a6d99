"""Mixed-model solve on the HIP GRG sweeps: the standardized kinship operator, batched
FP64 conjugate gradients, top principal components and the learned CG warm start."""

from pygrgl_spmv.lmm.cg import CgResult, conj_grad, shifted
from pygrgl_spmv.lmm.operator import StandardizedGrgOperator, center
from pygrgl_spmv.lmm.pcs import Eigenpairs, top_eigenpairs
from pygrgl_spmv.lmm.simulate import Phenotypes, simulate
from pygrgl_spmv.lmm.warmstart import Basis, WarmStartNet, closed_form_coefficients, features, start_from_coefficients, targets

__all__ = [
    "Basis", "CgResult", "Eigenpairs", "Phenotypes", "StandardizedGrgOperator", "WarmStartNet", "center",
    "closed_form_coefficients", "conj_grad", "features", "shifted", "simulate", "start_from_coefficients", "targets",
    "top_eigenpairs",
]
