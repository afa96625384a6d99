"""CPU oracle for the GRG matmul hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module, and only as the checker. The product
(``pygrgl_spmv.backends.hip``) never calls it.

It restates, with its own artifact reader, the reference's host arithmetic:

* level sweeps: ``pygrgl_spmv/grg/__init__.py:144-205`` (``_reference_run_up``,
  ``_reference_run_down``) and ``pygrgl_spmv/backends/reference.py:365-432``;
* init seeding: ``grg/__init__.py:119-141``; endpoint biases ``:208-253``;
* endpoint semantics of ``BoundGRG.matmul`` on a host runtime, ``:636-734``
  (by_individual expand/reduce, endpoint bias, emit_all_nodes permutation,
  missingness output);
* the band artifacts and their closed forms used by the reference's GPU
  exactness tests: ``pygrgl_spmv/tests/runtime/_streaming_cases.py:201-317``
  (``write_three_level_band_artifact``, ``write_overlap_band_artifact``) and
  ``:493-526`` (``expected_up`` / ``expected_down``).

Pinning (see ``tests/test_oracle.py``): the sweeps are checked against the band
closed forms (prefix-sum window sums, no sparse algebra), against a dense
brute-force path-count matrix on tiny DAGs, against the diploid X^T X closed
form for ``init='xtx'``, and by adjointness and linearity identities.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import scipy.sparse as sp

# --------------------------------------------------------------------------
# Independent .grg_spmv reader (keys of reference artifact.py:124-184).
# --------------------------------------------------------------------------


@dataclass
class OracleGRG:
    num_samples: int
    num_mutations: int
    num_nodes: int
    num_individuals: int
    ploidy: int
    level_offsets: np.ndarray
    node_perm: np.ndarray
    sample_to_individual: np.ndarray
    blocks: dict  # (d, s) -> csr (only non-empty)
    sel_mut: sp.csr_matrix
    sel_miss: sp.csr_matrix
    coal: np.ndarray | None
    vec_up_bias: np.ndarray
    vec_down_bias: np.ndarray
    xtx_up_bias: np.ndarray | None
    xtx_down_bias: np.ndarray | None

    @property
    def height(self) -> int:
        return int(self.level_offsets.size - 1)

    @property
    def num_edges(self) -> int:
        return int(sum(b.nnz for b in self.blocks.values()))


def load_artifact(path) -> OracleGRG:
    with np.load(Path(path), allow_pickle=False) as z:
        if str(z["grg_spmv_magic"]) != "grg_spmv" or int(z["grg_spmv_format_version"]) != 8:
            raise ValueError("not a version-8 .grg_spmv artifact")
        S, M, N = (int(z[k]) for k in ("num_samples", "num_mutations", "num_nodes"))
        off = np.asarray(z["level_offsets"], dtype=np.int64)
        blocks = {}
        for (d, s), shape in zip(np.asarray(z["scan_block_levels"]), np.asarray(z["scan_block_shapes"])):
            idx = np.asarray(z[f"A_blocks_{d}_{s}_indices"], dtype=np.int64)
            if idx.size == 0:
                continue
            ptr = np.asarray(z[f"A_blocks_{d}_{s}_indptr"], dtype=np.int64)
            blocks[(int(d), int(s))] = sp.csr_matrix((np.ones(idx.size), idx, ptr), shape=tuple(int(v) for v in shape))

        def sel(prefix):
            idx = np.asarray(z[f"{prefix}_indices"], dtype=np.int64)
            ptr = np.asarray(z[f"{prefix}_indptr"], dtype=np.int64)
            return sp.csr_matrix((np.ones(idx.size), idx, ptr), shape=(M, N))

        has_coal = bool(z["has_individual_coals"])
        has_xtx = bool(z["has_xtx_bias"])
        return OracleGRG(
            num_samples=S, num_mutations=M, num_nodes=N,
            num_individuals=int(z["num_individuals"]), ploidy=int(z["ploidy"]),
            level_offsets=off,
            node_perm=np.asarray(z["node_perm"], dtype=np.int64),
            sample_to_individual=np.asarray(z["sample_to_individual"], dtype=np.int64),
            blocks=blocks, sel_mut=sel("sel_mut"), sel_miss=sel("sel_miss"),
            coal=np.asarray(z["coalescence_counts"], dtype=np.int64) if has_coal else None,
            vec_up_bias=np.asarray(z["init_vector_up_bias"], dtype=np.float64),
            vec_down_bias=np.asarray(z["init_vector_down_bias"], dtype=np.float64),
            xtx_up_bias=np.asarray(z["init_xtx_up_bias"], dtype=np.float64) if has_xtx else None,
            xtx_down_bias=np.asarray(z["init_xtx_down_bias"], dtype=np.float64) if has_xtx else None,
        )


def from_state(state) -> OracleGRG:
    """Wrap an in-memory compiled state (fields of reference compile.py:18-46)."""
    blocks = {}
    for d, row in enumerate(state.A_blocks):
        for s, b in enumerate(row):
            if b.nnz:
                blocks[(d, s)] = sp.csr_matrix((np.ones(b.nnz), np.asarray(b.indices, dtype=np.int64), np.asarray(b.indptr, dtype=np.int64)), shape=b.shape)

    def f64sel(m):
        return sp.csr_matrix((np.ones(m.nnz), np.asarray(m.indices, dtype=np.int64), np.asarray(m.indptr, dtype=np.int64)), shape=m.shape)

    return OracleGRG(
        num_samples=int(state.num_samples), num_mutations=int(state.num_mutations), num_nodes=int(state.num_nodes),
        num_individuals=int(state.num_individuals), ploidy=int(state.ploidy),
        level_offsets=np.asarray(state.level_offsets, dtype=np.int64),
        node_perm=np.asarray(state.node_perm, dtype=np.int64),
        sample_to_individual=np.asarray(state.sample_to_individual, dtype=np.int64),
        blocks=blocks, sel_mut=f64sel(state.sel_mut), sel_miss=f64sel(state.sel_miss),
        coal=None if state.coalescence_counts is None else np.asarray(state.coalescence_counts, dtype=np.int64),
        vec_up_bias=np.asarray(state.init_vector_up_bias if state.init_vector_up_bias is not None else np.zeros(state.num_mutations), dtype=np.float64),
        vec_down_bias=np.asarray(state.init_vector_down_bias if state.init_vector_down_bias is not None else np.zeros(state.num_samples), dtype=np.float64),
        xtx_up_bias=None if state.init_xtx_up_bias is None else np.asarray(state.init_xtx_up_bias, dtype=np.float64),
        xtx_down_bias=None if state.init_xtx_down_bias is None else np.asarray(state.init_xtx_down_bias, dtype=np.float64),
    )


# --------------------------------------------------------------------------
# Sweeps (reference grg/__init__.py:119-205, reference.py:365-432).
# --------------------------------------------------------------------------


def _seed(g: OracleGRG, k: int, init_mode: str, init_payload, dtype) -> np.ndarray:
    v = np.zeros((g.num_nodes, k), dtype=dtype)
    if init_mode == "xtx":
        v += (2.0 * g.coal.astype(dtype))[:, None]
    elif init_mode == "vector":
        v += np.asarray(init_payload, dtype=dtype)[None, :]
    elif init_mode == "matrix":  # payload already in internal node order, (N, k)
        v += np.asarray(init_payload, dtype=dtype)
    elif init_mode != "none":
        raise ValueError(init_mode)
    return v


def sweep_up(g: OracleGRG, x_internal: np.ndarray, *, init_mode: str = "none", init_payload=None, dtype=np.float64) -> np.ndarray:
    """Node values after the up sweep; ``x_internal`` is (S, k)."""
    off = g.level_offsets
    v = _seed(g, x_internal.shape[1], init_mode, init_payload, dtype)
    v[: g.num_samples] += x_internal
    for d in range(1, g.height):
        for s in range(d):
            b = g.blocks.get((d, s))
            if b is not None:
                v[off[d] : off[d + 1]] += b @ v[off[s] : off[s + 1]]
    return v


def sweep_down(g: OracleGRG, y_internal: np.ndarray, *, miss_internal=None, init_mode: str = "none", init_payload=None, dtype=np.float64) -> np.ndarray:
    """Node values after the down sweep; ``y_internal`` is (M, k)."""
    off = g.level_offsets
    v = _seed(g, y_internal.shape[1], init_mode, init_payload, dtype)
    v += g.sel_mut.T @ y_internal
    if miss_internal is not None:
        v += g.sel_miss.T @ miss_internal
    for d in range(g.height - 2, -1, -1):
        for s in range(g.height - 1, d, -1):
            b = g.blocks.get((s, d))
            if b is not None:
                v[off[d] : off[d + 1]] += b.T @ v[off[s] : off[s + 1]]
    return v


def init_biases(g: OracleGRG):
    """(vec_up, vec_down, xtx_up, xtx_down) as reference grg/__init__.py:208-253 builds them."""
    one = np.ones(1)
    vu = g.sel_mut @ sweep_up(g, np.zeros((g.num_samples, 1)), init_mode="vector", init_payload=one)
    vd = sweep_down(g, np.zeros((g.num_mutations, 1)), init_mode="vector", init_payload=one)[: g.num_samples]
    xu = xd = None
    if g.coal is not None:
        xu = g.sel_mut @ sweep_up(g, np.zeros((g.num_samples, 1)), init_mode="xtx")
        xd = sweep_down(g, np.zeros((g.num_mutations, 1)), init_mode="xtx")[: g.num_samples]
        xu, xd = xu[:, 0], xd[:, 0]
    return vu[:, 0], vd[:, 0], xu, xd


def matmul(g: OracleGRG, x: np.ndarray, direction: str, *, by_individual: bool = False, init=None, miss=None,
           emit_all_nodes: bool = False, dtype=np.float64):
    """``BoundGRG.matmul`` on a host runtime (reference grg/__init__.py:636-734).

    ``x`` is (k, cols). ``init``: None, "xtx", a (k,) vector or a (k, N) matrix
    in original node order. For ``direction='up'`` with ``miss`` (k, M), the
    missingness output is accumulated into ``miss`` in place.
    Returns (k, out_cols).
    """
    k = x.shape[0]
    inv_perm = np.empty_like(g.node_perm)
    inv_perm[g.node_perm] = np.arange(g.node_perm.size)
    if init is None:
        mode, payload = "none", None
    elif isinstance(init, str):
        mode, payload = "xtx", None
    elif init.ndim == 1:
        mode, payload = "vector", np.asarray(init, dtype=dtype)
    else:
        mode, payload = "matrix", np.asarray(init, dtype=dtype)[:, g.node_perm].T  # internal order (N, k)
    if emit_all_nodes:
        b_mode, b_payload = mode, payload
    elif mode in ("vector", "xtx"):
        b_mode, b_payload = "none", None
    else:
        b_mode, b_payload = mode, payload
    xi = np.asarray(x, dtype=dtype).T
    if direction == "up":
        if by_individual:
            xi = xi[g.sample_to_individual]
        v = sweep_up(g, xi, init_mode=b_mode, init_payload=b_payload, dtype=dtype)
        if emit_all_nodes:
            return v[inv_perm].T.copy()
        out = np.asarray(g.sel_mut @ v, dtype=dtype)
        if mode == "xtx":
            out += g.xtx_up_bias[:, None]
        elif mode == "vector":
            out += g.vec_up_bias[:, None] * payload[None, :]
        if miss is not None:
            miss += np.asarray(g.sel_miss @ v, dtype=dtype).T
        return out.T.copy()
    miss_i = None if miss is None else np.asarray(miss, dtype=dtype).T
    v = sweep_down(g, xi, miss_internal=miss_i, init_mode=b_mode, init_payload=b_payload, dtype=dtype)
    if emit_all_nodes:
        return v[inv_perm].T.copy()
    out = v[: g.num_samples].copy()
    if mode == "xtx":
        out += g.xtx_down_bias[:, None]
    elif mode == "vector":
        out += g.vec_down_bias[:, None] * payload[None, :]
    if by_individual:
        acc = np.zeros((g.num_individuals, k), dtype=dtype)
        np.add.at(acc, g.sample_to_individual, out)
        out = acc
    return out.T.copy()


# --------------------------------------------------------------------------
# Independent checks that pin the sweeps.
# --------------------------------------------------------------------------


def dense_path_counts(g: OracleGRG) -> np.ndarray:
    """P[a, b] = number of directed paths from node a down to node b (a != b), P[a, a] = 1.

    Built from the full (N, N) parent->child adjacency by a fixed-point
    iteration P = I + A P, independent of the level order (tiny DAGs only).
    """
    n = g.num_nodes
    off = g.level_offsets
    A = np.zeros((n, n))
    for (d, s), b in g.blocks.items():
        A[off[d] : off[d + 1], off[s] : off[s + 1]] += b.toarray()
    P = np.eye(n)
    for _ in range(g.height + 1):
        P = np.eye(n) + A @ P
    return P


def dense_genotypes(g: OracleGRG) -> np.ndarray:
    """X (S, M): path counts from each mutation's node(s) down to each sample."""
    P = dense_path_counts(g)
    return (g.sel_mut.toarray() @ P[:, : g.num_samples]).T


def xtx_closed_form(g: OracleGRG) -> np.ndarray:
    """Per mutation, sum over individuals of (diploid genotype)^2 -- what up(ones, init='xtx') returns."""
    X = dense_genotypes(g)
    geno = np.zeros((g.num_individuals, g.num_mutations))
    np.add.at(geno, g.sample_to_individual, X)
    return (geno * geno).sum(axis=0)


def _window_sum(x: np.ndarray, *, shift: int, bandwidth: int) -> np.ndarray:
    """out[i] = sum_{j=0}^{bandwidth-1} x[(i + shift + j) mod n] via prefix sums (reference _streaming_cases.py:493-504)."""
    arr = np.asarray(x, dtype=np.float64)
    n = arr.shape[0]
    start = (np.arange(n) + int(shift)) % n
    pre = np.zeros((2 * n + 1, arr.shape[1]))
    np.cumsum(np.concatenate((arr, arr)), axis=0, out=pre[1:])
    return pre[start + int(bandwidth)] - pre[start]


def band_expected_up(x: np.ndarray, *, shifts, bandwidth: int) -> np.ndarray:
    """Closed form of the three-level band up sweep (reference _streaming_cases.py:515-517)."""
    level1 = _window_sum(x, shift=shifts[0], bandwidth=bandwidth)
    return _window_sum(x, shift=shifts[1], bandwidth=bandwidth) + _window_sum(level1, shift=shifts[2], bandwidth=bandwidth)


def band_expected_down(y: np.ndarray, *, shifts, bandwidth: int) -> np.ndarray:
    """Closed form of the three-level band down sweep (reference _streaming_cases.py:520-526)."""
    tshift = lambda s: -int(s) - int(bandwidth) + 1  # noqa: E731
    level1 = _window_sum(y, shift=tshift(shifts[2]), bandwidth=bandwidth)
    return _window_sum(level1, shift=tshift(shifts[0]), bandwidth=bandwidth) + _window_sum(y, shift=tshift(shifts[1]), bandwidth=bandwidth)


def band_block(n: int, bandwidth: int, shift: int, dtype=np.int32):
    """Row i holds columns (i + shift + j) mod n, j < bandwidth (reference _streaming_cases.py:201-212)."""
    cols = (np.arange(n)[:, None] + np.arange(bandwidth)[None, :] + int(shift)) % n
    indptr = np.arange(0, (n + 1) * bandwidth, bandwidth)
    return cols.reshape(-1).astype(dtype), indptr.astype(dtype)


def normwise_rel_err(actual: np.ndarray, expected: np.ndarray) -> float:
    """Per column (row of a (k, cols) result) max|a-b| / max|b|, maxed over columns."""
    a = np.atleast_2d(np.asarray(actual, dtype=np.float64))
    b = np.atleast_2d(np.asarray(expected, dtype=np.float64))
    num = np.abs(a - b).max(axis=1)
    den = np.maximum(np.abs(b).max(axis=1), np.finfo(np.float64).tiny)
    return float((num / den).max())


__all__ = [
    "OracleGRG",
    "band_block",
    "band_expected_down",
    "band_expected_up",
    "dense_genotypes",
    "dense_path_counts",
    "from_state",
    "init_biases",
    "load_artifact",
    "matmul",
    "normwise_rel_err",
    "sweep_down",
    "sweep_up",
    "xtx_closed_form",
]
