"""MI355X-native GRG sparse matmul (``pygrgl_spmv.backends.hip``).

The package root only re-exports CPU-safe symbols; GPU symbols live in
``pygrgl_spmv.backends.hip``.
"""

from pygrgl_spmv.grg import BoundGRG, RuntimeRequirements, convert

__all__ = ["BoundGRG", "RuntimeRequirements", "convert"]
