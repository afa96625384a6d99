"""Reader for the IGD genotype file among the reference's test data
(``pygrgl_spmv/tests/data/test-200-samples.miss.igd``) -- test infrastructure:
it is the independent check of the ``.grg`` reader, since both files hold the
same 200-individual dataset.

IGD (version 4) as this reader decodes it, little-endian: a 128-byte header
(u64 magic ``0x3A0C6FD7945A3481``, u64 version, u32 ploidy, u32 sparse
threshold, u64 variant rows, u64 individuals, u64 flags, u64 offsets of the
row index, the allele table, the individual ids and the variant ids); the row
index holds two u64 per row (bits 0-55 the position, bits 56-63 the row flags:
bit 0 sparse, bit 1 missing-data row; then the row's file offset). A sparse row
is a u32 count and that many u32 sample ids; a dense row is a bit vector over
the samples, most significant bit first. The allele table holds, per row, the
reference and the alternative allele, each a u32 length and its bytes; a
missing-data row has an empty alternative allele.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

IGD_MAGIC = 0x3A0C6FD7945A3481


@dataclass
class Igd:
    ploidy: int
    num_individuals: int
    positions: np.ndarray
    row_flags: np.ndarray
    rows: list          # sample ids per row
    refs: list
    alts: list

    @property
    def num_samples(self) -> int:
        return self.ploidy * self.num_individuals


def read_igd(path) -> Igd:
    b = open(path, "rb").read()
    magic, version, ploidy, _sparse, nv, ni, _flags, p_index, p_alleles, _p_ind, _p_var = struct.unpack_from("<QQIIQQQQQQQ", b, 0)
    assert magic == IGD_MAGIC and version == 4, (hex(magic), version)
    ns = ploidy * ni
    idx = np.frombuffer(b, dtype="<u8", count=2 * nv, offset=p_index).reshape(nv, 2)
    pos = idx[:, 0] & ((1 << 56) - 1)
    flags = (idx[:, 0] >> 56).astype(np.int64)
    rows = []
    for i in range(nv):
        at = int(idx[i, 1])
        if flags[i] & 1:
            c = struct.unpack_from("<I", b, at)[0]
            rows.append(np.frombuffer(b, dtype="<u4", count=c, offset=at + 4).astype(np.int64))
        else:
            bits = np.unpackbits(np.frombuffer(b, dtype=np.uint8, count=(ns + 7) // 8, offset=at), bitorder="big")[:ns]
            rows.append(np.flatnonzero(bits))
    refs, alts = [], []
    at = p_alleles
    for _ in range(nv):
        for out in (refs, alts):
            n = struct.unpack_from("<I", b, at)[0]
            out.append(b[at + 4:at + 4 + n].decode())
            at += 4 + n
    return Igd(ploidy=ploidy, num_individuals=ni, positions=pos, row_flags=flags, rows=rows, refs=refs, alts=alts)


def genotype_columns(igd: Igd, positions, refs, alts):
    """(X, Miss): samples x mutations 0/1 matrices for the given mutations, matched to
    IGD rows by (position, reference allele, alternative allele); Miss from the
    missing-data row at the mutation's position (empty if none)."""
    key = {}
    miss_at = {}
    for i in range(len(igd.rows)):
        if igd.row_flags[i] & 2:
            miss_at[int(igd.positions[i])] = igd.rows[i]
        else:
            key[(int(igd.positions[i]), igd.refs[i], igd.alts[i])] = i
    m = len(positions)
    X = np.zeros((igd.num_samples, m))
    Miss = np.zeros((igd.num_samples, m))
    for j in range(m):
        X[igd.rows[key[(int(positions[j]), refs[j], alts[j])]], j] = 1.0
        Miss[miss_at.get(int(positions[j]), np.zeros(0, np.int64)), j] = 1.0
    return X, Miss
