"""GRGL ``.grg`` ingestion on the CPU: the reader and ``convert`` against the
reference's own test data (``pygrgl_spmv/tests/data``, copied to
``tests/golden/grgl``).

Pins: the 200-sample GRG against the IGD genotype file of the same dataset
(every mutation's carriers and missing samples, through the oracle's sweeps of
the converted artifact); its coalescence counts against the X^T X diagonal of
those genotypes; both GRGs' coalescence counts against their definition
(individuals whose haplotypes both lie below the node and below no single
child); ``convert``'s path rules and errors (reference
``tests/runtime/test_convert.py``).
"""

from __future__ import annotations

import os
import shutil
from pathlib import Path, PurePath

import numpy as np
import pytest

from oracle import grg_oracle as O
from pygrgl_spmv import convert
from pygrgl_spmv.grg.artifact import artifact_path_for_grg, load_grg_spmv
from pygrgl_spmv.grg.grgl_file import decode_leb128, read_grg
from tests._igd import genotype_columns, read_igd

GOLDEN = Path(__file__).resolve().parent / "golden" / "grgl"
T200 = GOLDEN / "test-200-samples.miss.final.grg"
MSPRIME = GOLDEN / "msprime.example.igd.final.grg"
IGD = GOLDEN / "test-200-samples.miss.igd"


@pytest.fixture(scope="module")
def converted(tmp_path_factory):
    d = tmp_path_factory.mktemp("grgl")
    return {"t200": convert(T200, d, name="t200"), "msprime": convert(MSPRIME, d, name="msprime")}


def test_leb128_known_values():
    vals, starts = decode_leb128(np.array([0xAC, 0x02, 0x05, 0x80, 0x80, 0x01, 0x00], dtype=np.uint8))
    assert vals.tolist() == [300, 5, 16384, 0] and starts.tolist() == [0, 2, 3, 6]
    with pytest.raises(ValueError, match="truncated"):
        decode_leb128(np.array([0x05, 0x80], dtype=np.uint8))


def test_reader_header_and_sections():
    g = read_grg(T200)
    assert (g.version[0], g.ploidy, g.num_samples, g.num_mutations, g.num_nodes, g.num_edges) == (5, 2, 400, 10893, 12858, 116797)
    assert g.has_missing_data and g.coalescence_counts is not None
    assert g.individual_ids == [f"tsk_{i}" for i in range(200)]
    assert np.all(np.diff(g.positions.astype(np.int64)) >= 0)
    assert int(np.count_nonzero(g.mutation_node < 0)) == 2132          # variants with no carrier
    m = read_grg(MSPRIME)
    assert (m.num_samples, m.num_mutations, m.num_nodes, m.num_edges) == (20000, 4, 20004, 700)
    assert not m.has_missing_data and m.miss_node is None
    assert m.individual_ids[-1] == "tsk_9999"


def test_t200_genotypes_and_missingness_match_igd(converted):
    g = read_grg(T200)
    igd = read_igd(IGD)
    assert igd.num_samples == g.num_samples
    assert int(np.count_nonzero((igd.row_flags & 2) == 0)) == g.num_mutations
    X, Miss = genotype_columns(igd, g.positions, g.ref_alleles, g.alleles)
    o = O.load_artifact(converted["t200"])
    S = g.num_samples
    miss = np.zeros((S, g.num_mutations))
    got = O.matmul(o, np.eye(S), "up", miss=miss)
    np.testing.assert_array_equal(got, X)
    np.testing.assert_array_equal(miss, Miss)
    assert int(np.count_nonzero(Miss.any(axis=0))) == 18               # mutations with missing samples
    # down: X y, by individual
    y = np.random.default_rng(0).standard_normal((3, g.num_mutations))
    D = np.zeros((g.num_individuals, g.num_mutations))
    np.add.at(D, np.arange(S) // 2, X)
    np.testing.assert_allclose(O.matmul(o, y, "down", by_individual=True), y @ D.T, rtol=1e-12, atol=1e-10)


def test_t200_xtx_diagonal_matches_igd(converted):
    g = read_grg(T200)
    X, _ = genotype_columns(read_igd(IGD), g.positions, g.ref_alleles, g.alleles)
    D = np.zeros((g.num_individuals, g.num_mutations))
    np.add.at(D, np.arange(g.num_samples) // 2, X)
    o = O.load_artifact(converted["t200"])
    got = O.matmul(o, np.ones((1, g.num_samples)), "up", init="xtx")[0]
    np.testing.assert_array_equal(got, (D * D).sum(axis=0))


@pytest.mark.parametrize("path", [T200, MSPRIME], ids=["t200", "msprime"])
def test_coalescence_counts_match_their_definition(path):
    g = read_grg(path)
    S, P = g.num_samples, g.ploidy
    below = [None] * g.num_nodes
    full = np.zeros(g.num_nodes, dtype=np.int64)
    for v in range(g.num_nodes):
        if v < S:
            below[v] = np.array([v])
            continue
        kids = g.child_idx[g.child_ptr[v]:g.child_ptr[v + 1]]
        s = np.unique(np.concatenate([below[c] for c in kids])) if kids.size else np.zeros(0, np.int64)
        below[v] = s
        ind, cnt = np.unique(s // P, return_counts=True)
        full[v] = int(np.count_nonzero(cnt == P))
        expect = full[v] - sum(int(full[c]) for c in kids)
        assert g.coalescence_counts[v] == expect, (v, g.coalescence_counts[v], expect)


def test_converted_artifacts_load(converted):
    st = load_grg_spmv(converted["t200"], np.float64)
    assert st.has_missing_data and st.coalescence_counts is not None and st.num_individuals == 200
    assert st.sel_mut.nnz == 10893 - 2132 and st.sel_miss.nnz == 18
    st = load_grg_spmv(converted["msprime"], np.float64)
    assert st.num_samples == 20000 and not st.has_missing_data


# ---- convert(): the reference's path rules and errors ----

class _PathWrapper(os.PathLike):
    def __init__(self, value: str) -> None:
        self._value = value

    def __fspath__(self) -> str:
        return self._value


def test_convert_from_path_writes_artifact(tmp_path):
    artifact = convert(MSPRIME, tmp_path)
    assert artifact.exists() and artifact.suffix == ".grg_spmv"
    assert artifact == artifact_path_for_grg(MSPRIME, tmp_path)


def test_convert_avoids_basename_collisions(tmp_path):
    a, b = tmp_path / "a" / "sample.grg", tmp_path / "b" / "sample.grg"
    for p in (a, b):
        p.parent.mkdir(parents=True)
        shutil.copyfile(MSPRIME, p)
    out = tmp_path / "artifacts"
    ra, rb = convert(a, out), convert(b, out)
    assert ra == artifact_path_for_grg(a, out) and rb == artifact_path_for_grg(b, out) and ra != rb
    assert ra.exists() and rb.exists()


@pytest.mark.parametrize("factory", [PurePath, _PathWrapper], ids=["purepath", "pathlike-wrapper"])
def test_convert_accepts_generic_pathlikes(tmp_path, factory):
    artifact = convert(factory(str(MSPRIME)), tmp_path)
    assert artifact.exists() and artifact.suffix == ".grg_spmv"


def test_convert_object_source_requires_name(tmp_path):
    g = read_grg(MSPRIME)
    with pytest.raises(ValueError, match="name"):
        convert(g, tmp_path)
    assert convert(g, tmp_path, name="from-object") == tmp_path / "from-object.grg_spmv"


def test_convert_requires_output_dir_and_grg_suffix(tmp_path):
    with pytest.raises(ValueError, match="output_dir"):
        convert(MSPRIME, None)
    bad = tmp_path / "x.igd"
    shutil.copyfile(MSPRIME, bad)
    with pytest.raises(ValueError, match=r"\.grg"):
        convert(bad, tmp_path)


def _patched(tmp_path, edit):
    data = bytearray(MSPRIME.read_bytes())
    data = edit(data)
    p = tmp_path / "patched.grg"
    p.write_bytes(bytes(data))
    return p


@pytest.mark.parametrize("edit,match", [
    (lambda d: d[:8].__class__(b"\0" * 8) + d[8:], "magic"),
    (lambda d: d[:8] + bytes([4, 0]) + d[10:], "format 4"),
    (lambda d: d[:0x40] + bytes([3]) + d[0x41:], "flag"),
    (lambda d: d[:0x30] + bytes([1]) + d[0x31:], "population"),
    (lambda d: d + b"\0", "unread bytes"),
    (lambda d: d[:-5], "truncated"),
], ids=["magic", "version", "flags", "populations", "trailing", "truncated"])
def test_reader_rejects_what_it_does_not_know(tmp_path, edit, match):
    with pytest.raises(ValueError, match=match):
        read_grg(_patched(tmp_path, edit))
