"""Parity at the benchmark workloads' full sizes (GPU).

* 100,000 samples x 1,000,000 mutations (about 60 M edges), FP64, k = 1 and 8:
  the HIP up and down sweeps against the reference's MKL sweep arithmetic
  (``oracle/mkl_baseline.py``, restating ``mkl/backend.py:559-645``) within
  1e-12 normwise relative error per column, and exactly for all-ones input.
* 500,000 samples x 5,000,000 mutations with 5 % missing genotypes (about
  271 M edges): size-independent properties -- all-ones inputs give the MKL
  sweep's integer counts exactly, adjointness <y, X^T x> = <X y, x> holds to
  1e-12, and the FP32-compensated sweeps stay within 5e-7 of FP64 per column.

The artifacts are generated from fixed seeds into ``GRG_BENCH_DIR`` (shared
with ``bench.py``, so a later bench run reuses them).
"""

from __future__ import annotations

import os
from pathlib import Path

import numpy as np
import pytest

from oracle import grg_oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _artifact(name: str) -> Path:
    from pygrgl_spmv.synth import WORKLOADS, write_artifact

    cfg = WORKLOADS[name]
    d = Path(os.environ.get("GRG_BENCH_DIR", "/tmp/grg_bench"))
    d.mkdir(parents=True, exist_ok=True)
    p = d / f"{name}-seed{cfg.seed}.grg_spmv"
    if not p.exists():
        write_artifact(cfg, p)
    return p


def _runtime(path, dtype, k, use_miss):
    from pygrgl_spmv.backends.hip import HipPlanPair, HipRuntime, plan_hip_layout
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=use_miss, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=False)
    layout = plan_hip_layout(artifacts=[path], pair=HipPlanPair.default(), dtype=dtype, requirements=req,
                             vram_budget_bytes=0, ring_buffer_size=0, device=0, stream=0)
    return HipRuntime(layout)


def _col_err(got, want):
    return (np.linalg.norm(got - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-300)).max()


def test_c2_fp64_matches_mkl():
    from oracle.mkl_baseline import MklSweep, mkl_matmul

    path = _artifact("c2")
    g = O.load_artifact(path)
    print(f"c2: {g.num_samples} samples x {g.num_mutations} mutations, {g.num_edges} edges", flush=True)
    sweep = MklSweep(g, max_k=8)
    rng = np.random.default_rng(2026)
    try:
        with _runtime(path, np.float64, 8, False) as rt:
            (grg,) = rt.grgs
            for k in (1, 8):
                for direction, cols in (("up", g.num_samples), ("down", g.num_mutations)):
                    x = rng.standard_normal((k, cols))
                    assert _col_err(grg.matmul(x, direction), mkl_matmul(g, sweep, x, direction)) <= 1e-12
                    ones = np.ones((k, cols))
                    np.testing.assert_array_equal(grg.matmul(ones, direction), mkl_matmul(g, sweep, ones, direction))
    finally:
        sweep.close()


def test_c3_properties_at_full_size():
    from oracle.mkl_baseline import MklSweep, mkl_matmul

    path = _artifact("c3")
    g = O.load_artifact(path)
    print(f"c3: {g.num_samples} samples x {g.num_mutations} mutations, {g.num_edges} edges", flush=True)
    rng = np.random.default_rng(2027)
    k = 16
    x = rng.standard_normal((k, g.num_samples))
    y = rng.standard_normal((k, g.num_mutations))
    miss = rng.standard_normal((k, g.num_mutations))
    sweep = MklSweep(g, max_k=1)
    try:
        ones_up = mkl_matmul(g, sweep, np.ones((1, g.num_samples)), "up")
        ones_dn = mkl_matmul(g, sweep, np.ones((1, g.num_mutations)), "down", miss=np.ones((1, g.num_mutations)))
    finally:
        sweep.close()
    out = {}
    for dtype in (np.float64, np.float32):
        with _runtime(path, dtype, k, True) as rt:
            (grg,) = rt.grgs
            up = grg.matmul(x.astype(dtype), "up").astype(np.float64)
            dn = grg.matmul(y.astype(dtype), "down", miss=miss.astype(dtype)).astype(np.float64)
            out[np.dtype(dtype).name] = (up, dn)
            if dtype == np.float64:
                np.testing.assert_array_equal(grg.matmul(np.ones((1, g.num_samples)), "up"), ones_up)
                np.testing.assert_array_equal(
                    grg.matmul(np.ones((1, g.num_mutations)), "down", miss=np.ones((1, g.num_mutations))), ones_dn)
                # adjointness: <y, X^T x> = <X y, x> per column, relative to ||y|| ||X^T x||
                dn_y = grg.matmul(y, "down")
                lhs = np.einsum("km,km->k", y, up)
                rhs = np.einsum("ks,ks->k", dn_y, x)
                scale = np.linalg.norm(y, axis=1) * np.linalg.norm(up, axis=1)
                assert np.max(np.abs(lhs - rhs) / scale) <= 1e-12
    up64, dn64 = out["float64"]
    up32, dn32 = out["float32"]
    err = max(_col_err(up32, up64), _col_err(dn32, dn64))
    print(f"c3: FP32c max per-column relative error vs FP64 {err:.2e}", flush=True)
    assert err <= 5e-7
