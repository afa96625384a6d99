"""Build the native libraries in-tree.

* ``pygrgl_spmv/_lib/libgrg_hip.so``   -- HIP sweep kernels + C ABI (hipcc, gfx950)
* ``pygrgl_spmv/_lib/libgrg_mlp.so``   -- HIP warm-start network kernels (hipcc, gfx950)
* ``pygrgl_spmv/_lib/libgrg_synth.so`` -- synthetic GRG generator (g++)
* ``pygrgl_spmv/_lib/libgrg_host.so``  -- host-side node-order pass (g++)

Run ``python grg-spmv_amd/build.py`` (or ``__graft_entry__.build()``).
"""

from __future__ import annotations

import os
from pathlib import Path
import shutil
import subprocess
import sys

ROOT = Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
LIB = ROOT / "pygrgl_spmv" / "_lib"
INCLUDE = ROOT.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all HIP code (v_pk_* f32 forms).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops",)
    err = "\n".join(line for line in proc.stderr.splitlines() if not any(n in line for n in noise))
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{err}")
    if err.strip():
        print(err, file=sys.stderr)


def _atomic_build(cmd_for, target: Path) -> None:
    tmp = target.with_suffix(".so.tmp")
    _run(cmd_for(tmp))
    os.replace(tmp, target)


def build_hip() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    target = LIB / "libgrg_hip.so"
    _atomic_build(lambda out: [HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", "-o", str(out), str(CSRC / "grg_hip.hip")], target)
    return target


def build_synth() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    target = LIB / "libgrg_synth.so"
    cxx = shutil.which("g++") or "g++"
    _atomic_build(lambda out: [cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-Wextra",
                               f"-I{INCLUDE}", "-o", str(out), str(CSRC / "grg_synth.cpp")], target)
    return target


def build_host() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    target = LIB / "libgrg_host.so"
    cxx = shutil.which("g++") or "g++"
    _atomic_build(lambda out: [cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
                               f"-I{INCLUDE}", "-o", str(out), str(CSRC / "grg_host.cpp")], target)
    return target


def build_mlp() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    target = LIB / "libgrg_mlp.so"
    _atomic_build(lambda out: [HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", "-o", str(out), str(CSRC / "grg_mlp.hip")], target)
    return target


def build_all() -> list[Path]:
    return [build_hip(), build_mlp(), build_synth(), build_host()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
