"""Reader for GRGL's binary ``.grg`` files (format 5.x), without ``pygrgl``.

The reference loads ``.grg`` files through the third-party ``pygrgl`` (pinned
2.6, ``uv.lock:1278-1280``; ``pygrgl.load_immutable_grg`` at
``grg/__init__.py:760``) and compiles what it returns (``compile_grg``,
``compile.py:385-460``). ``pygrgl`` is not available here, so this module
decodes the file itself. The layout was read from the reference's own test data
(``pygrgl_spmv/tests/data/*.grg``); the 200-sample file is checked against the
IGD genotype file of the same dataset, mutation by mutation (carriers, missing
samples, positions, alleles), and the coalescence counts against the X^T X
diagonal of those genotypes.

Layout (little-endian), as far as this reader decodes it:

* bytes 0-127, header: u64 magic ``0xE9366C64DDC8C5B0``; u16 major version (5),
  u16 minor version, u16 ploidy, u16 zero; u64 samples, mutations, nodes, edges,
  populations; u64 not decoded (a range bound); u64 flags (bit 0: the
  coalescence-count section is present, bit 2: missing data); zero padding.
* u64 E, then E bytes of down edges: for every node with children, in node-id
  order, LEB128 varints holding the child count and then the child ids
  delta-coded (the first absolute, ascending).
* u64[nodes + 1]: byte offset of each node's list within those E bytes.
* u32[mutations]: the node of each mutation (``0x7FFFFFFF`` = none).
* 24-byte mutation records, sorted by position: 8 bytes of allele storage (byte
  0 is 0 and bytes 1-7 hold the reference allele then the alternative allele,
  NUL-padded), u64 position, 4 bytes not decoded, u32 length of the reference
  allele.
* (flag bit 0) u32[nodes - samples]: individual-coalescence count of every
  non-sample node.
* u64 B, then B bytes of individual ids (each a u8 length and its bytes), then
  u64[individuals + 1] offsets into them.
* (flag bit 2) u32[mutations]: the missingness node of each mutation.
* nothing after that: the reader requires the sections to end at the file's end.

Anything outside what these files exercise -- another major version, a
population table, out-of-line alleles longer than 7 bytes, other flag bits, a
layout without flag bit 0 -- is rejected with ``ValueError`` rather than
guessed. Mutation times are not decoded (they are written as NaN).
"""

from __future__ import annotations

import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

GRGL_MAGIC = 0xE9366C64DDC8C5B0
GRGL_MAJOR_VERSION = 5
INVALID_NODE = 0x7FFFFFFF
HEADER_BYTES = 128
FLAG_COALS = 1
FLAG_MISSING = 4
_KNOWN_FLAGS = FLAG_COALS | FLAG_MISSING
_RECORD = np.dtype([("alleles", "u1", (8,)), ("position", "<u8"), ("undecoded", "<u4"), ("ref_len", "<u4")])


@dataclass
class GrglGraph:
    """A decoded ``.grg``: the down-edge DAG as CSR over node ids, and the mutation table."""

    version: tuple
    ploidy: int
    num_samples: int
    num_mutations: int
    num_nodes: int
    num_edges: int
    flags: int
    child_ptr: np.ndarray          # int64, nodes + 1
    child_idx: np.ndarray          # int64, edges; ascending within a node, < the parent id
    mutation_node: np.ndarray      # int64, -1 = no node
    miss_node: np.ndarray | None   # int64, -1 = none; None without missing data
    positions: np.ndarray          # uint64
    ref_alleles: list
    alleles: list
    coalescence_counts: np.ndarray | None  # int64 per node (0 on samples)
    individual_ids: list

    @property
    def num_individuals(self) -> int:
        return self.num_samples // max(self.ploidy, 1)

    @property
    def has_missing_data(self) -> bool:
        return self.miss_node is not None and bool(np.any(self.miss_node >= 0))


class _Cursor:
    def __init__(self, data: bytes, path):
        self.data, self.pos, self.path = data, 0, path

    def take(self, n: int, what: str) -> memoryview:
        n = int(n)
        if n < 0 or self.pos + n > len(self.data):
            raise ValueError(f"{self.path}: truncated .grg file while reading {what}")
        view = memoryview(self.data)[self.pos:self.pos + n]
        self.pos += n
        return view

    def array(self, dtype, count: int, what: str) -> np.ndarray:
        dt = np.dtype(dtype)
        return np.frombuffer(self.take(dt.itemsize * int(count), what), dtype=dt, count=int(count))

    def u64(self, what: str) -> int:
        return int(self.array("<u8", 1, what)[0])


def decode_leb128(buf: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(values uint64, byte offset of each value's first byte) of a LEB128 byte stream."""
    buf = np.asarray(buf, dtype=np.uint8)
    if buf.size == 0:
        return np.zeros(0, np.uint64), np.zeros(0, np.int64)
    ends = np.flatnonzero(buf < 0x80)
    if ends.size == 0 or ends[-1] != buf.size - 1:
        raise ValueError("truncated LEB128 value at the end of the edge section")
    starts = np.empty_like(ends)
    starts[0] = 0
    starts[1:] = ends[:-1] + 1
    lengths = ends - starts + 1
    if int(lengths.max()) > 9:
        raise ValueError("LEB128 value longer than 9 bytes in the edge section")
    shift = (np.arange(buf.size, dtype=np.int64) - np.repeat(starts, lengths)) * 7
    vals = (buf & 0x7F).astype(np.uint64) << shift.astype(np.uint64)
    return np.add.reduceat(vals, starts), starts


def _decode_edges(edges: np.ndarray, offsets: np.ndarray, num_nodes: int, num_edges: int, path):
    values, starts = decode_leb128(edges)
    off = offsets.astype(np.int64)
    if off[0] != 0 or np.any(np.diff(off) < 0) or off[-1] != edges.size:
        raise ValueError(f"{path}: node offsets do not partition the edge section")
    first = np.searchsorted(starts, off)                      # varint index at each node's first byte
    has = off[:-1] < off[1:]
    fv = first[:-1][has]
    if np.any(fv >= starts.size) or np.any(starts[np.minimum(fv, max(starts.size - 1, 0))] != off[:-1][has]):
        raise ValueError(f"{path}: a node's edge list does not start on a varint boundary")
    nvals = np.diff(first)
    counts = np.zeros(num_nodes, dtype=np.int64)
    counts[has] = values[first[:-1][has]].astype(np.int64)
    if np.any(nvals[has] != counts[has] + 1) or np.any(nvals[~has] != 0):
        raise ValueError(f"{path}: a node's edge list length disagrees with its child count")
    if int(counts.sum()) != int(num_edges):
        raise ValueError(f"{path}: {int(counts.sum())} edges decoded, the header says {num_edges}")
    is_delta = np.ones(values.size, dtype=bool)
    is_delta[first[:-1][has]] = False
    deltas = values[is_delta].astype(np.int64)
    ptr = np.zeros(num_nodes + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    csum = np.cumsum(deltas)
    seg_first = ptr[:-1][counts > 0]
    base = np.repeat(csum[seg_first] - deltas[seg_first], counts[counts > 0])
    child = csum - base
    parent = np.repeat(np.arange(num_nodes, dtype=np.int64), counts)
    if child.size:
        if np.any(child >= parent) or np.any(child < 0):
            raise ValueError(f"{path}: a child id is not in [0, parent id)")
        later = np.ones(child.size, dtype=bool)
        later[seg_first] = False
        if np.any(deltas[later] == 0):
            raise ValueError(f"{path}: repeated child id in a node's edge list")
    return ptr, child


def _node_array(raw: np.ndarray, num_nodes: int, what: str, path) -> np.ndarray:
    v = raw.astype(np.int64)
    bad = (v != INVALID_NODE) & (v >= num_nodes)
    if np.any(bad):
        raise ValueError(f"{path}: {what} refers to node {int(v[bad][0])} of {num_nodes}")
    v[v == INVALID_NODE] = -1
    return v


def read_grg(path) -> GrglGraph:
    """Decode a GRGL ``.grg`` file (format 5.x) into a ``GrglGraph``."""
    path = Path(os.fspath(path))
    cur = _Cursor(path.read_bytes(), path)
    head = cur.take(HEADER_BYTES, "the header")
    magic = int(np.frombuffer(head, dtype="<u8", count=1)[0])
    if magic != GRGL_MAGIC:
        raise ValueError(f"{path}: not a GRGL .grg file (magic {magic:#x})")
    major, minor, ploidy, _ = (int(x) for x in np.frombuffer(head, dtype="<u2", count=4, offset=8))
    if major != GRGL_MAJOR_VERSION:
        raise ValueError(f"{path}: GRGL format {major}.{minor} is not supported (only {GRGL_MAJOR_VERSION}.x)")
    ns, nm, nn, ne, npop, _range, flags = (int(x) for x in np.frombuffer(head, dtype="<u8", count=7, offset=16))
    if npop != 0:
        raise ValueError(f"{path}: population tables are not supported ({npop} populations)")
    if flags & ~_KNOWN_FLAGS or not flags & FLAG_COALS:
        raise ValueError(f"{path}: unsupported GRGL flag bits {flags:#x}")
    if not (0 < ns <= nn < INVALID_NODE) or ploidy < 1 or ns % ploidy:
        raise ValueError(f"{path}: inconsistent header (samples {ns}, nodes {nn}, ploidy {ploidy})")

    edge_bytes = cur.u64("the edge section size")
    edges = np.frombuffer(cur.take(edge_bytes, "the edge section"), dtype=np.uint8)
    offsets = cur.array("<u8", nn + 1, "the node offsets")
    child_ptr, child_idx = _decode_edges(edges, offsets, nn, ne, path)
    if np.any(np.diff(child_ptr[:ns + 1]) != 0):
        raise ValueError(f"{path}: a sample node has children")

    mutation_node = _node_array(cur.array("<u4", nm, "the mutation nodes"), nn, "a mutation", path)
    recs = cur.array(_RECORD, nm, "the mutation records")
    if np.any(recs["alleles"][:, 0] != 0):
        raise ValueError(f"{path}: out-of-line (long) alleles are not supported")
    refs, alts = [], []
    for raw, rl in zip(recs["alleles"], recs["ref_len"]):
        text = bytes(raw[1:]).split(b"\0", 1)[0].decode("utf-8")
        if int(rl) > len(text):
            raise ValueError(f"{path}: reference allele length {int(rl)} exceeds the stored alleles {text!r}")
        refs.append(text[:int(rl)])
        alts.append(text[int(rl):])
    positions = recs["position"].copy()

    coal = None
    if flags & FLAG_COALS:
        coal = np.zeros(nn, dtype=np.int64)
        coal[ns:] = cur.array("<u4", nn - ns, "the coalescence counts").astype(np.int64)

    id_bytes = cur.u64("the individual id section size")
    id_blob = bytes(cur.take(id_bytes, "the individual ids"))
    ni = ns // ploidy
    id_off = cur.array("<u8", ni + 1, "the individual id offsets").astype(np.int64)
    if id_off[0] != 0 or id_off[-1] != id_bytes or np.any(np.diff(id_off) < 1):
        raise ValueError(f"{path}: individual id offsets do not partition the id section")
    ids = []
    for a, b in zip(id_off[:-1], id_off[1:]):
        n = id_blob[a]
        if a + 1 + n != b:
            raise ValueError(f"{path}: individual id length prefix disagrees with its offsets")
        ids.append(id_blob[a + 1:b].decode("utf-8"))

    miss = None
    if flags & FLAG_MISSING:
        miss = _node_array(cur.array("<u4", nm, "the missingness nodes"), nn, "a missingness entry", path)
    if cur.pos != len(cur.data):
        raise ValueError(f"{path}: {len(cur.data) - cur.pos} unread bytes after the last known section")
    return GrglGraph(version=(major, minor), ploidy=ploidy, num_samples=ns, num_mutations=nm, num_nodes=nn, num_edges=ne,
                     flags=flags, child_ptr=child_ptr, child_idx=child_idx, mutation_node=mutation_node, miss_node=miss,
                     positions=positions, ref_alleles=refs, alleles=alts, coalescence_counts=coal, individual_ids=ids)


def compile_grgl(graph: GrglGraph):
    """``CompiledOperatorState`` of a decoded GRG (the reference's ``compile_grg``,
    ``compile.py:385-460``: levels, stable height order, level blocks, selectors,
    coalescence counts, mutation table)."""
    from pygrgl_spmv.grg.compile import compile_dag

    return compile_dag(
        num_samples=graph.num_samples,
        child_ptr=graph.child_ptr,
        child_idx=graph.child_idx,
        mutation_node=graph.mutation_node,
        miss_node=graph.miss_node,
        ploidy=graph.ploidy,
        coalescence_counts=graph.coalescence_counts,
        mutation_positions=graph.positions.astype(np.float64),
        mutation_times=np.full(graph.num_mutations, np.nan),
        alleles=graph.alleles,
        ref_alleles=graph.ref_alleles,
    )


__all__ = ["GRGL_MAGIC", "GrglGraph", "INVALID_NODE", "compile_grgl", "decode_leb128", "read_grg"]
