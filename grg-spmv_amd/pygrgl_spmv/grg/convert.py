"""``convert``: compile a GRGL ``.grg`` into a ``.grg_spmv`` artifact.

Mirrors the reference's ``convert`` (``grg/__init__.py:737-787``): same
arguments, artifact path rule and errors. The reference loads the file through
``pygrgl``; here ``grgl_file.read_grg`` decodes it, and a decoded ``GrglGraph``
stands where the reference accepts an ``ImmutableGRG`` object.
"""

from __future__ import annotations

import os
from pathlib import Path

import numpy as np

from pygrgl_spmv.grg.artifact import artifact_path_for_grg, save_grg_spmv
from pygrgl_spmv.grg.grgl_file import GrglGraph, compile_grgl, read_grg


def convert(source, output_dir="pygrgl_spmv_artifacts", *, dtype=np.float64, name: str | None = None) -> Path:
    """Compile a GRG and write a ``.grg_spmv`` artifact.

    Path sources default to an artifact path derived from the resolved source path
    under ``output_dir``; ``name=`` writes ``output_dir / f"{name}.grg_spmv"`` instead.
    A ``GrglGraph`` source requires ``name``.
    """
    from pygrgl_spmv.grg import _build_init_biases

    dtype = np.dtype(dtype)
    if output_dir is None:
        raise ValueError("convert() requires an output_dir in the runtime-owned API")
    if isinstance(source, (str, os.PathLike)):
        source_path = Path(os.fspath(source))
        if source_path.suffix != ".grg":
            raise ValueError(f"expected a .grg file, got {source_path}")
        stem = source_path.stem if name is None else str(name)
        graph = read_grg(source_path)
    elif isinstance(source, GrglGraph):
        if name is None:
            raise ValueError("name is required when source is a decoded GRG object")
        stem = str(name)
        graph = source
    else:
        raise TypeError(f"convert() takes a .grg path or a GrglGraph, got {type(source).__name__}")

    compiled = compile_grgl(graph)
    del graph
    _build_init_biases(compiled, dtype)
    out_dir = Path(output_dir).expanduser()
    out_dir.mkdir(parents=True, exist_ok=True)
    if isinstance(source, (str, os.PathLike)) and name is None:
        artifact_path = artifact_path_for_grg(source_path, out_dir)
    else:
        artifact_path = out_dir / f"{stem}.grg_spmv"
    save_grg_spmv(compiled, artifact_path)
    return artifact_path


__all__ = ["convert"]
