"""Operator-surface semantics, helpers and the C ABI (CPU only).

Mirrors the reference's unit tests that need no GPU: requirement and spec
validation messages (``grg/__init__.py:63-84,454-516``), the op-surface
argument checks of ``matmul`` (``:564-596``, ``endtoend/test_matmul.py:169-181``),
sparse helpers (``test_sparse_utils.py``), the compile toys
(``test_compile_selectors.py:150-180``), level-pair order (``base.py:106-119``),
stream parsing (``_cuda_stream.py:108-174``) and the shared-library surface.
"""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from pygrgl_spmv.backends._hip_stream import HipStreamToken, parse_hip_stream
from pygrgl_spmv.backends.base import iter_direction_level_pairs, relink_stream_dependencies
from pygrgl_spmv.backends.types import Direction, InitMode, parse_direction, parse_init_mode, parse_sparse_format, parse_store
from pygrgl_spmv.grg import BoundGRG, RuntimeRequirements
from pygrgl_spmv.grg.compile import build_level_blocks, compile_dag, stable_height_order
from pygrgl_spmv.grg.sparse import binary_csr_from_parts, finalize_int_array, pick_small_signed_int_dtype
from tests._builders import full_requirements

REPO = Path(__file__).resolve().parents[1]


# ---- requirements / spec -------------------------------------------------


def test_runtime_requirements_validation():
    with pytest.raises(ValueError, match="max_k_up must be >= 1"):
        RuntimeRequirements(0, 1, False, False, False, False, False)
    with pytest.raises(ValueError, match="max_k_down must be >= 1"):
        RuntimeRequirements(1, 0, False, False, False, False, False)


class _FakeRuntime:
    def __init__(self, requirements):
        self.layout = SimpleNamespace(requirements=requirements, dtype=np.dtype(np.float64))

    def _prepare_matmul_cuda(self, grg, spec):
        return spec


def _grg(requirements=None, *, coal=True):
    state = SimpleNamespace(num_samples=4, num_individuals=2, num_mutations=3, num_nodes=7, coalescence_counts=np.zeros(7) if coal else None)
    return BoundGRG(_FakeRuntime(requirements or full_requirements(max_k_up=4, max_k_down=4)), 0, state, Path("x.grg_spmv"))


def test_prepare_spec_columns_and_init_split():
    g = _grg()
    spec = g.prepare_matmul_cuda(direction="up", k=2, by_individual=True, init_mode="vector")
    assert (spec.input_cols, spec.output_cols, spec.backend_init_mode, spec.apply_endpoint_bias) == (2, 3, InitMode.NONE, True)
    spec = g.prepare_matmul_cuda(direction="down", k=2, by_individual=True, init_mode="matrix")
    assert (spec.input_cols, spec.output_cols, spec.backend_init_mode, spec.apply_endpoint_bias) == (3, 2, InitMode.MATRIX, False)
    spec = g.prepare_matmul_cuda(direction="down", k=1, emit_all_nodes=True, init_mode="xtx")
    assert (spec.output_cols, spec.backend_init_mode, spec.apply_endpoint_bias) == (7, InitMode.XTX, False)


def test_prepare_spec_errors():
    g = _grg(full_requirements(max_k_up=2, max_k_down=2, need_up_miss_output=False, need_init_matrix=False))
    with pytest.raises(ValueError, match="k must be >= 1"):
        g.prepare_matmul_cuda(direction="up", k=0)
    with pytest.raises(RuntimeError, match="emit_all_nodes"):
        g.prepare_matmul_cuda(direction="up", k=1, emit_all_nodes=True, use_miss=True)
    with pytest.raises(ValueError, match="non-none init_mode"):
        g.prepare_matmul_cuda(direction="down", k=1, init_mode="vector", use_miss=True)
    with pytest.raises(ValueError, match="exceeds declared max_k_up=2"):
        g.prepare_matmul_cuda(direction="up", k=3)
    with pytest.raises(ValueError, match="UP miss output was not declared"):
        g.prepare_matmul_cuda(direction="up", k=1, use_miss=True)
    with pytest.raises(ValueError, match="init matrix use was not declared"):
        g.prepare_matmul_cuda(direction="up", k=1, init_mode="matrix")
    with pytest.raises(ValueError, match="coalescence counts"):
        _grg(coal=False).prepare_matmul_cuda(direction="up", k=1, init_mode="xtx")
    with pytest.raises(ValueError, match="Unknown init_mode"):
        g.prepare_matmul_cuda(direction="up", k=1, init_mode="bogus")
    with pytest.raises(ValueError, match="unknown direction|Unknown direction"):
        g.prepare_matmul_cuda(direction="sideways", k=1)


def test_matmul_argument_checks():
    g = _grg()
    with pytest.raises(TypeError, match="numpy.ndarray"):
        g.matmul([[1.0]], "up")
    with pytest.raises(ValueError, match="two-dimensional"):
        g.matmul(np.ones(4), "up")
    with pytest.raises(ValueError, match="non-zero"):
        g.matmul(np.ones((0, 4)), "up")
    with pytest.raises(ValueError, match="UP direction"):
        g.matmul(np.ones((1, 3)), "up")
    with pytest.raises(ValueError, match="DOWN direction"):
        g.matmul(np.ones((1, 4)), "down")
    with pytest.raises(RuntimeError, match="emit_all_nodes"):
        g.matmul(np.ones((1, 4)), "up", emit_all_nodes=True, miss=np.ones((1, 3)))
    with pytest.raises(ValueError, match="cannot be mixed"):
        g.matmul(np.ones((1, 4)), "up", init="xtx", miss=np.ones((1, 3)))
    # init parsing (endtoend/test_matmul.py:169-181)
    for bad in (np.ones((1, 200000)), np.ones((2, 4)), np.ones(10), "test"):
        with pytest.raises((ValueError, TypeError)):
            g._parse_init(bad, 1 if not isinstance(bad, np.ndarray) or bad.ndim == 2 else 9, np.dtype(np.float64), reorder_matrix=False)
    with pytest.raises(TypeError, match="dtype"):
        g._parse_init(np.ones(1, dtype=np.float32), 1, np.dtype(np.float64), reorder_matrix=False)


def test_parsers():
    assert parse_direction("UP") == Direction.UP and parse_direction(Direction.DOWN) == Direction.DOWN
    assert parse_init_mode("XTX") == InitMode.XTX
    assert parse_store("t").value == "T" and parse_sparse_format("csc").value == "CSC"
    assert str(Direction.UP) == "up" and Direction.UP == "up"
    for fn, bad in ((parse_direction, "x"), (parse_store, "Q"), (parse_sparse_format, "ell")):
        with pytest.raises(ValueError):
            fn(bad)


# ---- helpers ---------------------------------------------------------------


def test_iter_direction_level_pairs():
    assert list(iter_direction_level_pairs(Direction.UP, 3)) == [(1, 0, 0), (2, 0, 0), (2, 1, 1)]
    assert list(iter_direction_level_pairs(Direction.DOWN, 3)) == [(0, 2, 1), (0, 1, 0), (1, 2, 0)]
    with pytest.raises(ValueError):
        list(iter_direction_level_pairs(Direction.UP, -1))


def test_relink_stream_dependencies():
    from dataclasses import dataclass

    @dataclass(frozen=True)
    class Op:
        slot: int | None
        prev_in_slot: tuple | None = None

    ops = [[], [Op(0), Op(None)], [Op(1), Op(0)], [Op(0)]]
    up = relink_stream_dependencies([list(x) for x in ops], direction=Direction.UP)
    assert up[1][0].prev_in_slot is None and up[2][1].prev_in_slot == (1, 0) and up[3][0].prev_in_slot == (2, 1)
    down = relink_stream_dependencies([list(x) for x in ops], direction=Direction.DOWN)
    assert down[2][1].prev_in_slot is None and down[1][0].prev_in_slot == (2, 1)


def test_sparse_helpers():
    assert pick_small_signed_int_dtype(np.iinfo(np.int32).max) == np.int32
    assert pick_small_signed_int_dtype(np.iinfo(np.int32).max + 1) == np.int64
    arr = np.array([0, 1, 2], dtype=np.int32)
    assert finalize_int_array(arr, label="a") is arr
    assert finalize_int_array(np.array([0, 1], dtype=np.int64), label="a").dtype == np.int32
    assert finalize_int_array(np.empty(0, dtype=np.int64), label="a").dtype == np.int32
    with pytest.raises(TypeError, match="integer array"):
        finalize_int_array(np.array([0.5]), label="a")
    m = binary_csr_from_parts(indices=np.array([1, 0, 2], dtype=np.int32), indptr=np.array([0, 1, 3], dtype=np.int64), shape=(2, 3), shared_data=True)
    np.testing.assert_array_equal(m.toarray(), [[False, True, False], [True, False, True]])
    assert m.data.strides == (0,) and not m.data.flags.writeable
    with pytest.raises(ValueError, match="2 dimensions"):
        binary_csr_from_parts(indices=np.array([0], dtype=np.int32), indptr=np.array([0, 1], dtype=np.int32), shape=(1,))


def test_level_block_toy():
    """Reference test_compile_selectors.py:150-180: nodes 0,1 leaves; 2 -> 0; 3 -> 2."""
    child_ptr = np.array([0, 0, 0, 1, 2])
    child_idx = np.array([0, 2])
    levels = np.array([0, 0, 1, 2])
    perm, inv, off = stable_height_order(levels)
    blocks = build_level_blocks(child_ptr, child_idx, levels=levels, level_offsets=off, inv_node_perm=inv)
    np.testing.assert_array_equal(blocks[1][0].toarray(), [[True, False]])
    np.testing.assert_array_equal(blocks[2][0].toarray(), [[False, False]])
    np.testing.assert_array_equal(blocks[2][1].toarray(), [[True]])
    assert all(b.indices.dtype == np.int32 and b.indptr.dtype == np.int32 for row in blocks for b in row)


def test_compile_dag_orders_and_selectors():
    # samples 0,1,2; node 3 = {0,1}; node 4 = {3,2}; node 5 = {2} missing pool
    st = compile_dag(num_samples=3, child_ptr=np.array([0, 0, 0, 0, 2, 4, 5]), child_idx=np.array([0, 1, 3, 2, 2]),
                     mutation_node=np.array([3, 4, 1]), miss_node=np.array([5, -1, -1]))
    np.testing.assert_array_equal(st.level_offsets, [0, 3, 5, 6])
    np.testing.assert_array_equal(st.node_perm, [0, 1, 2, 3, 5, 4])
    np.testing.assert_array_equal(st.sel_mut.toarray().nonzero()[1], [3, 5, 1])
    assert st.sel_miss.nnz == 1 and st.has_missing_data
    with pytest.raises(ValueError, match="smaller than their parent"):
        compile_dag(num_samples=1, child_ptr=np.array([0, 1, 1]), child_idx=np.array([1]), mutation_node=np.array([0]))


# ---- streams ----------------------------------------------------------------


def test_parse_hip_stream():
    assert parse_hip_stream(0) == (0, None)
    tok = HipStreamToken(1234)
    assert parse_hip_stream(tok) == (1234, tok)
    assert parse_hip_stream(SimpleNamespace(__cuda_stream__=lambda: 77))[0] == 77
    assert parse_hip_stream(SimpleNamespace(cuda_stream=99))[0] == 99
    with pytest.raises(TypeError):
        parse_hip_stream(True)
    with pytest.raises(ValueError):
        parse_hip_stream(-1)
    with pytest.raises(ValueError, match="version"):
        parse_hip_stream(SimpleNamespace(__cuda_stream__=lambda: (1, 5)))
    with pytest.raises(TypeError):
        parse_hip_stream(SimpleNamespace(__cuda_stream__=lambda: (0, 1, 2)))


# ---- C ABI -----------------------------------------------------------------


def _declared_functions(header: Path) -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(grg_\w+)\s*\(", text, flags=re.M)


@pytest.mark.parametrize("header,lib", [("grg_hip.h", "libgrg_hip.so"), ("grg_synth.h", "libgrg_synth.so"),
                                        ("grg_host.h", "libgrg_host.so")])
def test_library_exports_every_declared_symbol(header, lib):
    path = REPO / "grg-spmv_amd" / "pygrgl_spmv" / "_lib" / lib
    if not path.exists():
        pytest.skip(f"{lib} not built")
    names = _declared_functions(REPO / "include" / header)
    assert names, "no declarations parsed"
    so = ctypes.CDLL(str(path))
    for name in names:
        assert hasattr(so, name), f"{lib} does not export {name}"


def test_ctypes_struct_layout_matches_c_header(tmp_path):
    """sizeof / offsetof of GrgHipGatherArgs from the C compiler equal ctypes' layout."""
    from pygrgl_spmv.backends.hip import ffi

    structs = {"GrgHipGatherArgs": ffi.GrgHipGatherArgs, "GrgHipSeg": ffi.GrgHipSeg, "GrgHipMap": ffi.GrgHipMap,
               "GrgHipJob": ffi.GrgHipJob}
    exprs, want = [], []
    for cname, cls in structs.items():
        exprs.append(f"sizeof({cname})")
        want.append(ctypes.sizeof(cls))
        for fname, _ in cls._fields_:
            exprs.append(f"offsetof({cname}, {fname})")
            want.append(getattr(cls, fname).offset)
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "grg_hip.h"\n'
        "int main(void){\n" + "".join(f'  printf("%zu\\n", (size_t)({e}));\n' for e in exprs) + "  return 0;}\n"
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", f"-I{REPO / 'include'}", "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert dict(zip(exprs, got)) == dict(zip(exprs, want))


def test_hip_library_reports_errors_without_gpu():
    from pygrgl_spmv.backends.hip import ffi

    if not ffi.LIB_PATH.exists():
        pytest.skip("libgrg_hip.so not built")
    lib = ffi.HipLibrary()
    bad = ffi.GrgHipGatherArgs()
    bad.dtype = 7
    with pytest.raises(RuntimeError, match="HIP grg_hip_gather failed: status 1: unknown dtype"):
        lib.gather(bad, 0)
    bad.dtype, bad.k, bad.cpl, bad.kl, bad.es, bad.row_end = ffi.F64, 3, 2, 1, 1, 1
    with pytest.raises(RuntimeError, match="multiple of cpl"):
        lib.gather(bad, 0)


def _light_args(rows, k, kl=4, es=4, cpl=1):
    from pygrgl_spmv.backends.hip import ffi

    a = ffi.GrgHipGatherArgs()
    a.dtype, a.k, a.cpl, a.kl, a.es, a.row_begin, a.row_end = ffi.F64, k, cpl, kl, es, 0, rows
    a.out, a.ld_out, a.heavy_threshold = 256, k, -1
    return a


def test_job_table_layout_without_gpu():
    """grg_hip_jobs_build: jobs laid end to end within each phase (light blocks of
    4 waves x 64 / (kl * es) rows, times the column chunks), phase offsets, errors."""
    from pygrgl_spmv.backends.hip import ffi

    if not ffi.LIB_PATH.exists():
        pytest.skip("libgrg_hip.so not built")
    lib = ffi.HipLibrary()
    args = [_light_args(100, 8, kl=4, es=4), _light_args(33, 16, kl=4, es=16, cpl=2), _light_args(5, 1, kl=1, es=64)]
    jobs, off, blocks = lib.jobs_build(args, [0, 0, 2])
    # job 0: 16 rows per block -> 7 blocks x ceil(8 / 4) = 2 chunks; job 1: 4 rows per block -> 9 x 2;
    # job 2: 4 rows per block -> 2 x 1
    assert [jobs[i].gx for i in range(3)] == [7, 9, 2]
    assert [jobs[i].gy for i in range(3)] == [2, 2, 1]
    assert [jobs[i].block_begin for i in range(3)] == [0, 14, 0]
    assert list(off) == [0, 2, 2, 3] and list(blocks) == [32, 0, 2]
    with pytest.raises(RuntimeError, match="non-decreasing"):
        lib.jobs_build(args, [1, 0, 2])
    tiled = _light_args(10, 4)
    tiled.flags = ffi.FLAG_TILED
    with pytest.raises(RuntimeError, match="job 0"):
        lib.jobs_build([tiled], [0])
