"""Artifact and layout builders shared by the tests.

The band artifacts restate the reference's synthetic streaming fixtures
(``pygrgl_spmv/tests/runtime/_streaming_cases.py:201-360``): three (or four)
levels of n nodes, each level block a circulant band of ``bandwidth`` ones per
row, samples = level 0, mutations = the top level through an identity
selector. Their sweeps have closed forms (``oracle.grg_oracle.band_expected_*``).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle.grg_oracle import band_block
from pygrgl_spmv.grg import RuntimeRequirements, _build_init_biases
from pygrgl_spmv.grg.artifact import save_grg_spmv
from pygrgl_spmv.grg.compile import CompiledOperatorState, _encode_alleles
from pygrgl_spmv.grg.sparse import binary_csr_from_parts

DEFAULT_GPU_BUDGET = 1_000_000_000


def full_requirements(*, max_k_up=8, max_k_down=8, need_down_miss_input=True, need_up_miss_output=True,
                      need_init_vector=True, need_init_matrix=True, need_init_xtx=True) -> RuntimeRequirements:
    return RuntimeRequirements(
        max_k_up=int(max_k_up), max_k_down=int(max_k_down),
        need_down_miss_input=bool(need_down_miss_input), need_up_miss_output=bool(need_up_miss_output),
        need_init_vector=bool(need_init_vector), need_init_matrix=bool(need_init_matrix), need_init_xtx=bool(need_init_xtx),
    )


def hip_pair(*, up: bool = True, down: bool = True):
    from pygrgl_spmv.backends.hip import HipPlanPair

    return HipPlanPair.default(up=up, down=down)


def build_hip_layout(artifacts, *, dtype=np.float64, requirements=None, pair=None, vram_budget_bytes=DEFAULT_GPU_BUDGET,
                     ring_buffer_size=0, allow_residency=True, device=0, stream=0):
    from pygrgl_spmv.backends.hip import plan_hip_layout

    return plan_hip_layout(
        artifacts=[Path(p) for p in artifacts], pair=hip_pair() if pair is None else pair, dtype=np.dtype(dtype),
        requirements=full_requirements() if requirements is None else requirements,
        vram_budget_bytes=int(vram_budget_bytes), ring_buffer_size=int(ring_buffer_size),
        allow_residency=bool(allow_residency), device=device, stream=stream,
    )


def synthetic_state(*, blocks, level_offsets, num_samples, num_mutations, num_nodes, sel_mut, sel_miss, dtype=np.float64) -> CompiledOperatorState:
    """Identity-permutation state with biases (reference _streaming_cases.py:231-271)."""
    sdt = np.dtype(np.asarray(level_offsets).dtype)
    perm = np.arange(num_nodes, dtype=sdt)
    alleles, offs = _encode_alleles(["A"] * num_mutations)
    state = CompiledOperatorState(
        A_blocks=blocks, level_offsets=np.asarray(level_offsets), node_perm=perm, inv_node_perm=perm.copy(),
        sel_mut=sel_mut, sel_miss=sel_miss, num_samples=int(num_samples), num_mutations=int(num_mutations),
        num_nodes=int(num_nodes), ploidy=1, num_individuals=int(num_samples),
        num_edges=int(sum(b.nnz for row in blocks for b in row)), has_missing_data=False,
        sample_to_individual=np.arange(num_samples, dtype=sdt), mutation_positions=np.arange(num_mutations, dtype=np.float64),
        mutation_times=np.zeros(num_mutations), mutation_alleles=alleles, mutation_allele_offsets=offs,
        mutation_ref_alleles=alleles.copy(), mutation_ref_allele_offsets=offs.copy(), coalescence_counts=None,
    )
    _build_init_biases(state, np.dtype(dtype))
    return state


def _band_csr(n, bw, shift, dtype):
    """Binary CSR that keeps the given structural dtype (scipy would narrow int64 to int32)."""
    import scipy.sparse as sp

    idx, ptr = band_block(n, bw, shift, dtype)
    m = sp.csr_matrix((n, n), dtype=np.bool_)
    m.data = np.ones(idx.size, dtype=np.bool_)
    m.indices = idx
    m.indptr = ptr
    return m


def write_band_artifact(artifact_dir, name, *, n, bandwidth, levels=3, struct_dtype=np.int32, shifts=None) -> Path:
    """3-level (blocks (1,0),(2,0),(2,1)) or 4-level band artifact, shifts k*bandwidth by default."""
    n, bw = int(n), int(bandwidth)
    sdt = np.dtype(struct_dtype)
    pairs = [(d, s) for d in range(1, levels) for s in range(d)]
    shifts = tuple(i * bw for i in range(len(pairs))) if shifts is None else tuple(int(v) for v in shifts)
    path = Path(artifact_dir) / f"{name}.grg_spmv"
    if path.exists():
        return path
    blocks = [[] for _ in range(levels)]
    for (d, s), sh in zip(pairs, shifts):
        blocks[d].append(_band_csr(n, bw, sh, sdt))
    top = (levels - 1) * n
    sel_mut = binary_csr_from_parts(indices=np.arange(top, top + n, dtype=sdt), indptr=np.arange(n + 1, dtype=sdt), shape=(n, levels * n), shared_data=True)
    sel_miss = binary_csr_from_parts(indices=np.empty(0, dtype=sdt), indptr=np.zeros(n + 1, dtype=sdt), shape=(n, levels * n), shared_data=True)
    state = synthetic_state(blocks=blocks, level_offsets=np.arange(0, (levels + 1) * n, n, dtype=sdt),
                            num_samples=n, num_mutations=n, num_nodes=levels * n, sel_mut=sel_mut, sel_miss=sel_miss)
    save_grg_spmv(state, path)
    return path


def write_synthetic_artifact(artifact_dir, name, cfg) -> Path:
    from pygrgl_spmv.synth import write_artifact

    path = Path(artifact_dir) / f"{name}.grg_spmv"
    if not path.exists():
        write_artifact(cfg, path)
    return path


def bench_dir() -> Path:
    """Where the benchmark workloads' artifacts are cached (shared with ``bench.py``)."""
    import os

    d = Path(os.environ.get("GRG_BENCH_DIR", "/tmp/grg_bench"))
    d.mkdir(parents=True, exist_ok=True)
    return d


def workload_artifact(name: str) -> Path:
    """The named synthetic workload (``pygrgl_spmv.synth.WORKLOADS``), generated once from its seed."""
    from pygrgl_spmv.synth import WORKLOADS, write_artifact

    cfg = WORKLOADS[name]
    p = bench_dir() / f"{name}-seed{cfg.seed}.grg_spmv"
    if not p.exists():
        write_artifact(cfg, p)
    return p


def chromosome_artifacts(count: int = 22, samples: int = 200_000, workers: int = 8) -> list[Path]:
    """The per-chromosome workload (``synth.chromosome_workloads``), generated in parallel processes."""
    from concurrent.futures import ProcessPoolExecutor

    from pygrgl_spmv.synth import chromosome_workloads, write_artifact

    d = bench_dir()
    cfgs = chromosome_workloads(count, samples)
    paths = [d / f"chr{i + 1}-{samples}-seed{c.seed}.grg_spmv" for i, c in enumerate(cfgs)]
    todo = [(c, p) for c, p in zip(cfgs, paths) if not p.exists()]
    if todo:
        with ProcessPoolExecutor(max(1, min(int(workers), len(todo)))) as ex:
            list(ex.map(write_artifact, [c for c, _ in todo], [p for _, p in todo]))
    return paths
