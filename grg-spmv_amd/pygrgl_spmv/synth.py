"""Seeded synthetic GRG workloads written as ``.grg_spmv`` artifacts.

``pygrgl`` (which the reference uses to compile real GRGs) is not part of this
package, so workloads come from the native generator in ``libgrg_synth.so``
(model described in ``csrc/grg_synth.cpp``). The output is a regular
``CompiledOperatorState``; endpoint biases are filled by ``_build_init_biases``
exactly as the reference's ``convert()`` does (``grg/__init__.py:776``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace
from pathlib import Path

import numpy as np

from pygrgl_spmv.grg import _build_init_biases
from pygrgl_spmv.grg.artifact import save_grg_spmv
from pygrgl_spmv.grg.compile import CompiledOperatorState, _encode_alleles, _invert_permutation
from pygrgl_spmv.grg.sparse import binary_csr_from_parts, finalize_int_array

_LIB_DIR = Path(__file__).resolve().parent / "_lib"


class _Params(ctypes.Structure):
    _fields_ = [
        ("num_samples", ctypes.c_int64),
        ("num_mutations", ctypes.c_int64),
        ("seed", ctypes.c_uint64),
        ("ploidy", ctypes.c_int32),
        ("num_threads", ctypes.c_int32),
        ("num_trees", ctypes.c_int64),
        ("num_segments", ctypes.c_int64),
        ("keep_fraction", ctypes.c_double),
        ("window_fraction", ctypes.c_double),
        ("missing_fraction", ctypes.c_double),
        ("missing_pool", ctypes.c_int64),
    ]


class _Result(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in ("num_samples", "num_mutations", "num_nodes", "num_edges", "num_levels", "ploidy")] + [
        ("level_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("node_perm", ctypes.POINTER(ctypes.c_int64)),
        ("block_edge_offset", ctypes.POINTER(ctypes.c_int64)),
        ("block_indptr_offset", ctypes.POINTER(ctypes.c_int64)),
        ("block_indices", ctypes.POINTER(ctypes.c_int32)),
        ("block_indptr", ctypes.POINTER(ctypes.c_int32)),
        ("mutation_node", ctypes.POINTER(ctypes.c_int64)),
        ("miss_node", ctypes.POINTER(ctypes.c_int64)),
        ("coalescence_counts", ctypes.POINTER(ctypes.c_int64)),
        ("mutation_positions", ctypes.POINTER(ctypes.c_double)),
        ("mutation_times", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        path = _LIB_DIR / "libgrg_synth.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build() (or grg-spmv_amd/build.py)")
        lib = ctypes.CDLL(str(path))
        lib.grg_synth_generate.argtypes = [ctypes.POINTER(_Params), ctypes.POINTER(_Result)]
        lib.grg_synth_generate.restype = ctypes.c_int
        lib.grg_synth_free.argtypes = [ctypes.POINTER(_Result)]
        lib.grg_synth_free.restype = None
        lib.grg_synth_last_error.argtypes = []
        lib.grg_synth_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


@dataclass(frozen=True)
class SynthConfig:
    """Generator parameters; ``num_trees``/``keep_fraction``/``window_fraction`` shape the DAG."""

    num_samples: int
    num_mutations: int
    seed: int
    ploidy: int = 1
    num_trees: int = 100
    num_segments: int = 16
    keep_fraction: float = 0.1
    window_fraction: float = 0.05
    missing_fraction: float = 0.0
    missing_pool: int = 64
    num_threads: int = 0


def _copy(ptr, n: int, dtype) -> np.ndarray:
    if n == 0:
        return np.empty(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).astype(dtype, copy=True)


def generate_state(cfg: SynthConfig, *, with_biases: bool = True, bias_dtype=np.float64) -> CompiledOperatorState:
    """Run the generator and wrap its output as a ``CompiledOperatorState``."""
    lib = _load()
    p = _Params(
        num_samples=int(cfg.num_samples), num_mutations=int(cfg.num_mutations), seed=int(cfg.seed) & (2**64 - 1),
        ploidy=int(cfg.ploidy), num_threads=int(cfg.num_threads), num_trees=int(cfg.num_trees),
        num_segments=int(cfg.num_segments), keep_fraction=float(cfg.keep_fraction),
        window_fraction=float(cfg.window_fraction), missing_fraction=float(cfg.missing_fraction),
        missing_pool=int(cfg.missing_pool),
    )
    r = _Result()
    if lib.grg_synth_generate(ctypes.byref(p), ctypes.byref(r)) != 0:
        raise RuntimeError(f"grg_synth_generate failed: {lib.grg_synth_last_error().decode()}")
    try:
        S, M, N, E, L = int(r.num_samples), int(r.num_mutations), int(r.num_nodes), int(r.num_edges), int(r.num_levels)
        off = _copy(r.level_offsets, L + 1, np.int64)
        node_perm = finalize_int_array(_copy(r.node_perm, N, np.int64), label="node_perm")
        eoff = _copy(r.block_edge_offset, L * L + 1, np.int64)
        poff = _copy(r.block_indptr_offset, L * L + 1, np.int64)
        indices = _copy(r.block_indices, E, np.int32)
        indptr = _copy(r.block_indptr, int(poff[-1]), np.int32)
        mut_node = _copy(r.mutation_node, M, np.int64)
        miss_node = _copy(r.miss_node, M, np.int64)
        coal = _copy(r.coalescence_counts, N, np.int64) if bool(r.coalescence_counts) else None
        pos = _copy(r.mutation_positions, M, np.float64)
        times = _copy(r.mutation_times, M, np.float64)
    finally:
        lib.grg_synth_free(ctypes.byref(r))
    sizes = np.diff(off)
    blocks = []
    for d in range(L):
        row = []
        for s in range(d):
            b = d * L + s
            row.append(
                binary_csr_from_parts(
                    indices=indices[eoff[b] : eoff[b + 1]],
                    indptr=indptr[poff[b] : poff[b + 1]],
                    shape=(int(sizes[d]), int(sizes[s])),
                    shared_data=True,
                )
            )
        blocks.append(row)

    def selector(nodes: np.ndarray):
        has = nodes >= 0
        ptr = np.zeros(M + 1, dtype=np.int64)
        np.cumsum(has, out=ptr[1:])
        return binary_csr_from_parts(
            indices=finalize_int_array(nodes[has], label="selector.indices"),
            indptr=finalize_int_array(ptr, label="selector.indptr"),
            shape=(M, N),
        )

    ploidy = int(cfg.ploidy)
    alleles, allele_off = _encode_alleles(["A"] * M)
    refs, ref_off = _encode_alleles(["G"] * M)
    state = CompiledOperatorState(
        A_blocks=blocks,
        level_offsets=finalize_int_array(off, label="level_offsets"),
        node_perm=node_perm,
        inv_node_perm=_invert_permutation(node_perm),
        sel_mut=selector(mut_node),
        sel_miss=selector(miss_node),
        num_samples=S,
        num_mutations=M,
        num_nodes=N,
        ploidy=ploidy,
        num_individuals=S // ploidy,
        num_edges=E,
        has_missing_data=bool(np.any(miss_node >= 0)),
        sample_to_individual=finalize_int_array(np.arange(S, dtype=np.int64) // ploidy, label="sample_to_individual"),
        mutation_positions=pos,
        mutation_times=times,
        mutation_alleles=alleles,
        mutation_allele_offsets=allele_off,
        mutation_ref_alleles=refs,
        mutation_ref_allele_offsets=ref_off,
        coalescence_counts=coal,
    )
    if with_biases:
        _build_init_biases(state, np.dtype(bias_dtype))
    return state


def write_artifact(cfg: SynthConfig, path, *, bias_dtype=np.float64) -> Path:
    """Generate ``cfg`` and save it as a ``.grg_spmv`` artifact at ``path``."""
    state = generate_state(cfg, with_biases=True, bias_dtype=bias_dtype)
    out = Path(path)
    save_grg_spmv(state, out)
    return out


def dag_stats(state: CompiledOperatorState) -> dict:
    return {
        "samples": int(state.num_samples),
        "mutations": int(state.num_mutations),
        "nodes": int(state.num_nodes),
        "edges": int(state.num_edges),
        "levels": int(len(state.level_offsets) - 1),
        "sel_mut_nnz": int(state.sel_mut.nnz),
        "sel_miss_nnz": int(state.sel_miss.nnz),
    }


# Named workloads (sizes and seeds of the benchmark configurations).
# C1: 2,000 x 20,000 haploid; C2: 100k x 1M; C3: 500k x 5M diploid with a 5 %
# missingness pool and coalescence counts; C4: per-chromosome GRGs over 200k
# samples; C5: 200k x 2M diploid.
WORKLOADS: dict[str, SynthConfig] = {
    "c1": SynthConfig(num_samples=2_000, num_mutations=20_000, seed=12345, num_trees=1600, keep_fraction=0.35, window_fraction=0.10, num_segments=8),
    "c2": SynthConfig(num_samples=100_000, num_mutations=1_000_000, seed=12346, num_trees=4000, keep_fraction=0.14, window_fraction=0.12, num_segments=64),
    "c3": SynthConfig(num_samples=500_000, num_mutations=5_000_000, seed=12347, ploidy=2, num_trees=4000, keep_fraction=0.14, window_fraction=0.12, num_segments=64, missing_fraction=0.05, missing_pool=64),
    "c5": SynthConfig(num_samples=200_000, num_mutations=2_000_000, seed=12349, ploidy=2, num_trees=4000, keep_fraction=0.14, window_fraction=0.12, num_segments=64),
}


def chromosome_workloads(num_chromosomes: int = 22, num_samples: int = 200_000, seed: int = 12348) -> list[SynthConfig]:
    """Per-chromosome GRGs with mutation counts drawn from U[20k, 400k]."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(20_000, 400_001, size=num_chromosomes)
    base = WORKLOADS["c5"]
    return [
        replace(base, num_samples=num_samples, num_mutations=int(m), seed=seed + i,
                num_trees=max(20, int(m) // 3000), num_segments=16)
        for i, m in enumerate(sizes)
    ]


__all__ = ["SynthConfig", "WORKLOADS", "chromosome_workloads", "dag_stats", "generate_state", "write_artifact"]
