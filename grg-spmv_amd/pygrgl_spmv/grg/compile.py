"""Compiled operator state and DAG compilation into level blocks.

``CompiledOperatorState`` has the fields of the reference's
``pygrgl_spmv/grg/compile.py:18-46``. The reference builds it from a ``pygrgl``
GRG (``compile_grg``, ``compile.py:385-460``); ``pygrgl`` is a third-party
native module that is not part of this package, so here the same layout is
built from a plain child-list DAG (``compile_dag``) with the same invariants:

* node level = 0 for childless nodes, else 1 + max child level
  (``compile.py:56-64``);
* stable height order: nodes sorted by level, ties kept in node-id order, and
  the samples ``0..S-1`` stay the prefix of level 0 (``compile.py:67-90,370-382``);
* one binary CSR block per (parent level, child level) pair with level-local,
  sorted, unique column indices (``compile.py:93-220``);
* mutation and missingness selectors as M x N binary CSR over internal node
  ids (``compile.py:223-296``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from pygrgl_spmv.grg.sparse import binary_csr_from_parts, finalize_int_array


@dataclass
class CompiledOperatorState:
    """Compiled operator state shared between artifact I/O and runtime setup."""

    A_blocks: list | None
    level_offsets: np.ndarray
    node_perm: np.ndarray
    inv_node_perm: np.ndarray
    sel_mut: sp.csr_matrix
    sel_miss: sp.csr_matrix
    num_samples: int
    num_mutations: int
    num_nodes: int
    ploidy: int
    num_individuals: int
    num_edges: int
    has_missing_data: bool
    sample_to_individual: np.ndarray
    mutation_positions: np.ndarray
    mutation_times: np.ndarray
    mutation_alleles: np.ndarray
    mutation_allele_offsets: np.ndarray
    mutation_ref_alleles: np.ndarray
    mutation_ref_allele_offsets: np.ndarray
    coalescence_counts: np.ndarray | None
    init_vector_up_bias: np.ndarray | None = None
    init_vector_down_bias: np.ndarray | None = None
    init_xtx_up_bias: np.ndarray | None = None
    init_xtx_down_bias: np.ndarray | None = None


def _invert_permutation(perm: np.ndarray) -> np.ndarray:
    """Inverse of a dense permutation, in the permutation's narrowed dtype."""
    p = finalize_int_array(perm, label="perm")
    inv = np.empty(p.size, dtype=p.dtype)
    inv[p] = np.arange(p.size, dtype=p.dtype)
    return inv


def node_levels_from_children(num_nodes: int, child_ptr: np.ndarray, child_idx: np.ndarray) -> np.ndarray:
    """Level of every node: 0 if childless, else 1 + max child level.

    Requires children to have smaller node ids than their parents (a
    topological numbering, as GRGs use); vectorised over a frontier sweep.
    """
    n = int(num_nodes)
    ptr = np.asarray(child_ptr, dtype=np.int64)
    idx = np.asarray(child_idx, dtype=np.int64)
    deg = np.diff(ptr)
    parent_of_edge = np.repeat(np.arange(n, dtype=np.int64), deg)
    if idx.size and np.any(idx >= parent_of_edge):
        raise ValueError("child node ids must be smaller than their parent's id")
    levels = np.zeros(n, dtype=np.int32)
    # Longest-path levels by repeated relaxation; converges in (height) rounds.
    while idx.size:
        cand = np.zeros(n, dtype=np.int32)
        np.maximum.at(cand, parent_of_edge, levels[idx] + 1)
        if np.array_equal(cand, levels):
            break
        levels = cand
    return levels


def stable_height_order(levels: np.ndarray):
    """(node_perm, inv_node_perm, level_offsets): internal = nodes sorted by level, stable."""
    lv = np.asarray(levels, dtype=np.int64)
    height = 0 if lv.size == 0 else int(lv.max()) + 1
    counts = np.bincount(lv, minlength=height)
    offsets = np.zeros(height + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    node_perm = np.argsort(lv, kind="stable").astype(np.int64)
    node_perm = finalize_int_array(node_perm, label="node_perm")
    return node_perm, _invert_permutation(node_perm), finalize_int_array(offsets, label="level_offsets")


def build_level_blocks(child_ptr, child_idx, *, levels, level_offsets, inv_node_perm) -> list:
    """Per (parent level d, child level s<d) binary CSR with level-local sorted columns."""
    ptr = np.asarray(child_ptr, dtype=np.int64)
    idx = np.asarray(child_idx, dtype=np.int64)
    lv = np.asarray(levels, dtype=np.int64)
    off = np.asarray(level_offsets, dtype=np.int64)
    inv = np.asarray(inv_node_perm, dtype=np.int64)
    height = off.size - 1
    n = lv.size
    parent = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
    p_level = lv[parent]
    c_level = lv[idx]
    if idx.size and np.any(c_level >= p_level):
        raise RuntimeError("invalid DAG: a child level is not below its parent level")
    p_row = inv[parent] - off[p_level]
    c_col = inv[idx] - off[c_level]
    order = np.lexsort((c_col, p_row, c_level, p_level))
    p_level, c_level, p_row, c_col = p_level[order], c_level[order], p_row[order], c_col[order]
    if c_col.size > 1:
        dup = (np.diff(p_level) == 0) & (np.diff(c_level) == 0) & (np.diff(p_row) == 0) & (np.diff(c_col) == 0)
        if np.any(dup):
            raise RuntimeError("duplicate DAG edge detected")
    key = p_level * max(height, 1) + c_level
    blocks: list[list[sp.csr_matrix]] = []
    for d in range(height):
        rows_d = int(off[d + 1] - off[d])
        row_blocks = []
        for s in range(d):
            cols_s = int(off[s + 1] - off[s])
            k = d * max(height, 1) + s
            a = int(np.searchsorted(key, k, side="left"))
            b = int(np.searchsorted(key, k, side="right"))
            if a == b:
                row_blocks.append(
                    binary_csr_from_parts(
                        indices=np.empty(0, dtype=np.int32),
                        indptr=np.zeros(rows_d + 1, dtype=np.int32),
                        shape=(rows_d, cols_s),
                    )
                )
                continue
            counts = np.bincount(p_row[a:b], minlength=rows_d)
            indptr = np.zeros(rows_d + 1, dtype=np.int64)
            np.cumsum(counts, out=indptr[1:])
            row_blocks.append(
                binary_csr_from_parts(
                    indices=finalize_int_array(c_col[a:b], label=f"A_blocks[{d}][{s}].indices"),
                    indptr=finalize_int_array(indptr, label=f"A_blocks[{d}][{s}].indptr"),
                    shape=(rows_d, cols_s),
                    shared_data=True,
                )
            )
        blocks.append(row_blocks)
    return blocks


def build_selector(node_of_mutation: np.ndarray, *, inv_node_perm: np.ndarray, num_nodes: int) -> sp.csr_matrix:
    """M x N selector with one entry per mutation whose node id is >= 0 (-1 = no node)."""
    nodes = np.asarray(node_of_mutation, dtype=np.int64)
    has = nodes >= 0
    indptr = np.zeros(nodes.size + 1, dtype=np.int64)
    np.cumsum(has, out=indptr[1:])
    indices = np.asarray(inv_node_perm, dtype=np.int64)[nodes[has]]
    return binary_csr_from_parts(
        indices=finalize_int_array(indices, label="selector.indices"),
        indptr=finalize_int_array(indptr, label="selector.indptr"),
        shape=(nodes.size, int(num_nodes)),
    )


def _encode_alleles(alleles) -> tuple[np.ndarray, np.ndarray]:
    """UTF-8 byte buffer + uint32 offsets (reference ``compile.py:326-343``)."""
    encoded = [str(a).encode("utf-8") for a in alleles]
    offsets = np.zeros(len(encoded) + 1, dtype=np.uint64)
    np.cumsum([len(e) for e in encoded], out=offsets[1:])
    if offsets[-1] > np.iinfo(np.uint32).max:
        raise ValueError("allele offset overflow: total byte count exceeds uint32 max")
    return np.frombuffer(b"".join(encoded), dtype=np.uint8).copy(), offsets.astype(np.uint32)


def compile_dag(
    *,
    num_samples: int,
    child_ptr: np.ndarray,
    child_idx: np.ndarray,
    mutation_node: np.ndarray,
    miss_node: np.ndarray | None = None,
    ploidy: int = 1,
    coalescence_counts: np.ndarray | None = None,
    mutation_positions: np.ndarray | None = None,
    mutation_times: np.ndarray | None = None,
    alleles=None,
    ref_alleles=None,
) -> CompiledOperatorState:
    """Compile a child-list DAG (node ids topologically numbered, samples first).

    ``coalescence_counts`` (per original node id, int64) enables ``init='xtx'``.
    """
    S = int(num_samples)
    ptr = np.asarray(child_ptr, dtype=np.int64)
    N = int(ptr.size - 1)
    if S <= 0 or N <= 0:
        raise ValueError("compile_dag() only supports non-empty DAGs with at least one sample and one node.")
    levels = node_levels_from_children(N, ptr, child_idx)
    if np.any(levels[:S] != 0):
        raise RuntimeError("sample nodes must be childless (level 0)")
    node_perm, inv_node_perm, level_offsets = stable_height_order(levels)
    if not np.array_equal(np.asarray(node_perm[:S]), np.arange(S)):
        raise RuntimeError("Stable height order must keep sample nodes as an internal prefix")
    blocks = build_level_blocks(ptr, child_idx, levels=levels, level_offsets=level_offsets, inv_node_perm=inv_node_perm)
    M = int(np.asarray(mutation_node).size)
    sel_mut = build_selector(mutation_node, inv_node_perm=inv_node_perm, num_nodes=N)
    miss = np.full(M, -1, dtype=np.int64) if miss_node is None else np.asarray(miss_node, dtype=np.int64)
    sel_miss = build_selector(miss, inv_node_perm=inv_node_perm, num_nodes=N)
    ploidy = max(int(ploidy), 1)
    s2i = finalize_int_array(np.arange(S, dtype=np.int64) // ploidy, label="sample_to_individual")
    coal = None
    if coalescence_counts is not None:
        coal = np.asarray(coalescence_counts, dtype=np.int64)[np.asarray(node_perm, dtype=np.int64)]
    alleles, allele_off = _encode_alleles(["A"] * M if alleles is None else alleles)
    refs, ref_off = _encode_alleles(["G"] * M if ref_alleles is None else ref_alleles)
    return CompiledOperatorState(
        A_blocks=blocks,
        level_offsets=level_offsets,
        node_perm=node_perm,
        inv_node_perm=inv_node_perm,
        sel_mut=sel_mut,
        sel_miss=sel_miss,
        num_samples=S,
        num_mutations=M,
        num_nodes=N,
        ploidy=ploidy,
        num_individuals=(S + ploidy - 1) // ploidy,
        num_edges=int(np.asarray(child_idx).size),
        has_missing_data=bool(sel_miss.nnz > 0),
        sample_to_individual=s2i,
        mutation_positions=np.arange(M, dtype=np.float64) if mutation_positions is None else np.asarray(mutation_positions, dtype=np.float64),
        mutation_times=np.zeros(M, dtype=np.float64) if mutation_times is None else np.asarray(mutation_times, dtype=np.float64),
        mutation_alleles=alleles,
        mutation_allele_offsets=allele_off,
        mutation_ref_alleles=refs,
        mutation_ref_allele_offsets=ref_off,
        coalescence_counts=coal,
    )


__all__ = ["CompiledOperatorState", "_invert_permutation", "compile_dag"]
