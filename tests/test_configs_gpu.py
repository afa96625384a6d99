"""The benchmark configurations through the HIP path, each at its stated size (GPU).

* C1, 2,000 samples x 20,000 mutations, k = 1: up and down sweeps against the
  CPU oracle (the ReferenceRuntime's SciPy arithmetic), all-ones inputs exactly
  and N(0, 1) inputs within 1e-12 (FP64), FP32c beside it.
* C3, 500,000 x 5,000,000 with 5 % missing genotypes, k = 32 and 64, with the
  XTX init reserved (``need_init_xtx=True``): integer inputs equal to the MKL
  sweep (the MklRuntime arithmetic, ``mkl/backend.py:559-645``) exactly, N(0, 1)
  inputs within 1e-12; the XTX term: the artifact's stored endpoint biases equal
  the MKL sweep of 2 * coalescence counts (reference ``grg/__init__.py:208-253``),
  and ``up(x, init='xtx') - up(x)`` (likewise down) equals them exactly (a
  non-emit XTX init is the endpoint bias alone, ``grg/__init__.py:498-502``);
  FP32c per-column error against the MKL sweep, reported and bounded.
* C4, 22 per-chromosome GRGs over 200,000 samples with k from 1 to 64: under a
  budget that streams most blocks through a ring of 4 slots, every GRG's eager
  up and down ops against the oracle on two of their columns, and all 44 sweeps
  as one grouped job-table batch (streamed blocks between its phases) against
  the per-GRG ops; under the 32 GB
  budget (every block resident), the job-table forms (persistent and grouped,
  eager, captured and replayed) against the per-GRG ops.
* C5, 200,000 x 2,000,000: the standardized kinship operator with FP32c sweeps,
  CG to a 1e-6 residual on 8 phenotypes, and the same solves with FP64 sweeps.
  The FP32c solutions' true residuals under the FP64 operator, and the distance
  between the two solutions, are checked against the bounds that the residuals
  and the shift delta give (H - delta I is positive semi-definite).

Parity of every kernel variant on small GRGs lives in ``test_hip_parity.py`` and
``test_hip_runtime.py``; the runtimes here use the default launch shapes
(``autotune=False``) so the suite stays within minutes.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import grg_oracle as O
from tests._builders import chromosome_artifacts, workload_artifact

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

FP64_TOL = 1e-12
FP32_TOL = 5e-7


def _runtime(paths, dtype, *, k, miss=False, xtx=False, budget=0, ring=0, autotune=False):
    from pygrgl_spmv.backends.hip import HipPlanPair, HipRuntime, plan_hip_layout
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=miss, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=xtx)
    layout = plan_hip_layout(artifacts=list(paths), pair=HipPlanPair.default(), dtype=dtype, requirements=req,
                             vram_budget_bytes=int(budget), ring_buffer_size=int(ring), device=0, stream=0)
    return HipRuntime(layout, autotune=autotune)


def _col_err(got, want):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    return np.linalg.norm(got - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-300)


# ---------------------------------------------------------------------------
# C1
# ---------------------------------------------------------------------------


def test_c1_against_oracle():
    path = workload_artifact("c1")
    g = O.load_artifact(path)
    print(f"c1: {g.num_samples} samples x {g.num_mutations} mutations, {g.num_nodes} nodes, {g.num_edges} edges, "
          f"{g.height} levels", flush=True)
    rng = np.random.default_rng(2026)
    inputs = {d: (np.ones((1, c)), rng.standard_normal((1, c))) for d, c in (("up", g.num_samples), ("down", g.num_mutations))}
    want = {d: (O.matmul(g, a, d), O.matmul(g, b, d)) for d, (a, b) in inputs.items()}
    for dtype, tol in ((np.float64, FP64_TOL), (np.float32, FP32_TOL)):
        with _runtime([path], dtype, k=1, autotune=True) as rt:
            (grg,) = rt.grgs
            for d, (ones, x) in inputs.items():
                np.testing.assert_array_equal(np.asarray(grg.matmul(ones.astype(dtype), d), np.float64), want[d][0])
                err = O.normwise_rel_err(grg.matmul(x.astype(dtype), d), want[d][1])
                print(f"c1 {np.dtype(dtype).name} {d}: rel err {err:.2e}", flush=True)
                assert err <= tol


# ---------------------------------------------------------------------------
# C3 at k = 32 and 64, XTX init
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def c3_reference():
    """The C3 artifact, its oracle form, and MKL-sweep results for integer and N(0, 1)
    inputs at k = 64 (the k = 32 ops take the first 32 columns)."""
    from oracle.mkl_baseline import MklSweep, mkl_matmul

    path = workload_artifact("c3")
    g = O.load_artifact(path)
    print(f"c3: {g.num_samples} x {g.num_mutations}, {g.num_edges} edges", flush=True)
    rng = np.random.default_rng(3264)
    S, M, k = g.num_samples, g.num_mutations, 64
    ins = {"x_int": rng.integers(-2, 3, size=(k, S)).astype(np.float64),
           "y_int": rng.integers(-2, 3, size=(k, M)).astype(np.float64),
           "m_int": rng.integers(-2, 3, size=(k, M)).astype(np.float64),
           "x": rng.standard_normal((k, S)), "y": rng.standard_normal((k, M)), "m": rng.standard_normal((k, M))}
    sweep = MklSweep(g, max_k=k)
    try:
        out = {"up_int": mkl_matmul(g, sweep, ins["x_int"], "up"),
               "dn_int": mkl_matmul(g, sweep, ins["y_int"], "down", miss=ins["m_int"]),
               "up": mkl_matmul(g, sweep, ins["x"], "up"),
               "dn": mkl_matmul(g, sweep, ins["y"], "down", miss=ins["m"])}
        # the XTX terms: the sweeps of 2 * coalescence counts (what the stored biases hold)
        v = (2.0 * g.coal.astype(np.float64))[:, None].copy()
        sweep.sweep(v, "up")
        out["xtx_up"] = np.asarray(g.sel_mut @ v)[:, 0]
        v = (2.0 * g.coal.astype(np.float64))[:, None].copy()
        sweep.sweep(v, "down")
        out["xtx_dn"] = v[:S, 0].copy()
    finally:
        sweep.close()
    print("c3: MKL reference sweeps done", flush=True)
    return path, g, ins, out


def test_c3_wide_k_fp64_with_xtx(c3_reference):
    path, g, ins, want = c3_reference
    np.testing.assert_array_equal(np.asarray(g.xtx_up_bias, np.float64), want["xtx_up"])
    np.testing.assert_array_equal(np.asarray(g.xtx_down_bias, np.float64), want["xtx_dn"])
    with _runtime([path], np.float64, k=64, miss=True, xtx=True) as rt:
        (grg,) = rt.grgs
        for k in (32, 64):
            got = grg.matmul(ins["x_int"][:k], "up")
            np.testing.assert_array_equal(got, want["up_int"][:k])
            got_x = grg.matmul(ins["x_int"][:k], "up", init="xtx")
            np.testing.assert_array_equal(got_x - got, np.broadcast_to(want["xtx_up"], (k, g.num_mutations)))
            got = grg.matmul(ins["y_int"][:k], "down", miss=ins["m_int"][:k])
            np.testing.assert_array_equal(got, want["dn_int"][:k])
            plain = grg.matmul(ins["y_int"][:k], "down")
            got_x = grg.matmul(ins["y_int"][:k], "down", init="xtx")
            np.testing.assert_array_equal(got_x - plain, np.broadcast_to(want["xtx_dn"], (k, g.num_samples)))
            e_up = _col_err(grg.matmul(ins["x"][:k], "up"), want["up"][:k]).max()
            e_dn = _col_err(grg.matmul(ins["y"][:k], "down", miss=ins["m"][:k]), want["dn"][:k]).max()
            print(f"c3 FP64 k={k}: integer inputs and XTX terms exact; N(0,1) max column rel err up {e_up:.2e}, "
                  f"down {e_dn:.2e}", flush=True)
            assert e_up <= FP64_TOL and e_dn <= FP64_TOL


def test_c3_wide_k_fp32c_error_against_fp64_oracle(c3_reference):
    path, g, ins, want = c3_reference
    with _runtime([path], np.float32, k=64, miss=True, xtx=True) as rt:
        (grg,) = rt.grgs
        for k in (32, 64):
            # node values of these integer inputs stay below 2^24 in magnitude, so FP32 storage
            # and the compensated sums keep them exact
            np.testing.assert_array_equal(np.asarray(grg.matmul(ins["x_int"][:k].astype(np.float32), "up"), np.float64),
                                          want["up_int"][:k])
            np.testing.assert_array_equal(
                np.asarray(grg.matmul(ins["y_int"][:k].astype(np.float32), "down", miss=ins["m_int"][:k].astype(np.float32)),
                           np.float64), want["dn_int"][:k])
            e_up = _col_err(grg.matmul(ins["x"][:k].astype(np.float32), "up"), want["up"][:k])
            e_dn = _col_err(grg.matmul(ins["y"][:k].astype(np.float32), "down", miss=ins["m"][:k].astype(np.float32)),
                            want["dn"][:k])
            print(f"c3 FP32c k={k}: per-column rel err vs the FP64 oracle: up max {e_up.max():.2e} median "
                  f"{np.median(e_up):.2e}; down max {e_dn.max():.2e} median {np.median(e_dn):.2e}", flush=True)
            assert e_up.max() <= FP32_TOL and e_dn.max() <= FP32_TOL


# ---------------------------------------------------------------------------
# C4: 22 per-chromosome GRGs
# ---------------------------------------------------------------------------




def _c4_streamed_budget(paths) -> int:
    """A budget an eighth of the way from the all-streamed layout's bytes to full residency:
    the planner keeps the largest blocks resident and streams the rest through the ring."""
    from pygrgl_spmv.chromosomes import streaming_budget

    return streaming_budget(paths, ring=4)


def test_c4_streamed_ring_per_grg_ops_against_oracle():
    from pygrgl_spmv.chromosomes import mixed_k

    paths = chromosome_artifacts()
    budget = _c4_streamed_budget(paths)
    with _runtime(paths, np.float64, k=64, budget=budget, ring=4) as rt:
        lay = rt.layout
        blocks = [b for a in lay.artifacts for b in (*a.blocks_up, *a.blocks_down)]
        streamed = [b for b in blocks if not b.resident]
        print(f"c4 streamed layout under {budget / 1e9:.2f} GB: {len(streamed)} of {len(blocks)} blocks streamed through "
              f"{lay.allocated_ring_buffer_size} slots ({sum(b.nbytes for b in streamed) / 1e9:.2f} GB), "
              f"largest streamed block {max(b.nbytes for b in streamed) / 1e6:.1f} MB", flush=True)
        assert streamed and len(streamed) < len(blocks) and lay.allocated_ring_buffer_size == 4
        rng = np.random.default_rng(4400)
        for i, (grg, path) in enumerate(zip(rt.grgs, paths)):
            k = mixed_k(i)
            g = O.load_artifact(path)
            cols = sorted({0, k - 1})
            x = rng.choice(np.array([-1.0, 1.0]), size=(k, g.num_samples))
            np.testing.assert_array_equal(grg.matmul(x, "up")[cols], O.matmul(g, x[cols], "up"))
            y = rng.standard_normal((k, g.num_mutations))
            err = O.normwise_rel_err(grg.matmul(y, "down")[cols], O.matmul(g, y[cols], "down"))
            assert err <= FP64_TOL, (i, err)
            y = rng.choice(np.array([-1.0, 1.0]), size=(k, g.num_mutations))
            np.testing.assert_array_equal(grg.matmul(y, "down")[cols], O.matmul(g, y[cols], "down"))
        print(f"c4 streamed: 22 GRGs, up exact on +-1 inputs, down exact on +-1 and within {FP64_TOL:g} on N(0,1)",
              flush=True)


def test_c4_streamed_batched_grouped_against_per_grg_ops():
    """Under the same streaming budget and ring of 4 slots, the 22 GRGs' up and down sweeps
    (mixed k) as ONE grouped job-table batch, its streamed blocks copied and accumulated
    between phases: every output equal to the per-GRG op's on +-1 inputs, and within 1e-12
    on N(0, 1) inputs."""
    import torch

    from pygrgl_spmv.chromosomes import mixed_k

    paths = chromosome_artifacts()
    budget = _c4_streamed_budget(paths)
    with _runtime(paths, np.float64, k=64, budget=budget, ring=4) as rt:
        assert any(not b.resident for a in rt.layout.artifacts for b in (*a.blocks_up, *a.blocks_down))
        ks = [mixed_k(i) for i in range(len(paths))]
        entries = [dict(grg=g, direction=d, k=k) for g, k in zip(rt.grgs, ks) for d in ("up", "down")]
        batch = rt.prepare_batched_sweep(entries, persistent=False)
        assert not batch.capturable
        rng = np.random.default_rng(4410)
        for kind in ("pm1", "normal"):
            inputs = []
            for e, view in zip(entries, batch.inputs):
                cols = e["grg"].num_samples if e["direction"] == "up" else e["grg"].num_mutations
                x = (rng.choice(np.array([-1.0, 1.0]), size=(e["k"], cols)) if kind == "pm1"
                     else rng.standard_normal((e["k"], cols)))
                view.copy_(torch.from_numpy(x))
                inputs.append(x)
            batch()
            batch.check()
            got = [out.cpu().numpy() for out in batch.outputs]
            worst = 0.0
            for e, x, out in zip(entries, inputs, got):
                want = e["grg"].matmul(x, e["direction"])
                if kind == "pm1":
                    np.testing.assert_array_equal(out, want)
                else:
                    worst = max(worst, O.normwise_rel_err(out, want))
            assert worst <= FP64_TOL, worst
        print(f"c4 streamed batch: {len(entries)} sweeps in one grouped job table, {batch.num_phases} phases, "
              f"{len(batch._streamed)} streamed block copies per run; exact on +-1, max rel err {worst:.2e}", flush=True)


def test_c4_resident_batched_forms_against_per_grg_ops():
    """The configuration's 32 GB budget keeps every block resident: every op captured in one
    graph and replayed bit for bit; the persistent and grouped job-table forms (up and down
    sweeps as two batches, or all 44 sweeps as one) within 1e-12 of the per-GRG ops, captured
    and replayed bit for bit (``chromosomes.run`` raises otherwise)."""
    from pygrgl_spmv.chromosomes import run

    res = run(chromosome_artifacts(), budget_bytes=32_000_000_000, ring=4, replays=2, autotune=False)
    print({k: v for k, v in res.items() if k not in ("edges", "ks")}, flush=True)
    assert res["streamed_blocks"] == 0 and res["captured"] and res["graph_replays_checked"] == 2
    for mode in ("persistent", "grouped", "persistent_one_batch", "grouped_one_batch"):
        assert res[f"batched_{mode}"]["max_rel_err_vs_per_grg_ops"] <= FP64_TOL


# ---------------------------------------------------------------------------
# C5: CG on the standardized kinship operator
# ---------------------------------------------------------------------------


def test_c5_cg_fp32c_solutions_against_fp64():
    import torch

    from pygrgl_spmv.lmm.cg import conj_grad, shifted
    from pygrgl_spmv.lmm.operator import StandardizedGrgOperator, center
    from pygrgl_spmv.lmm.simulate import simulate

    path = workload_artifact("c5")
    k = 8
    with _runtime([path], np.float64, k=k, xtx=True) as rt64, _runtime([path], np.float32, k=k, xtx=True) as rt32:
        op64 = StandardizedGrgOperator(rt64, k=k)
        op32 = StandardizedGrgOperator(rt32, k=k)
        print(f"c5: {op64.n} individuals, {op64.m_proj} model variants", flush=True)
        phen = simulate(op64, k, seed=5150)
        res32 = conj_grad(shifted(op32, phen.delta), phen.y, rel_tol=1e-6, max_iter=2000)
        res64 = conj_grad(shifted(op64, phen.delta), phen.y, rel_tol=1e-6, max_iter=2000)
        print(f"c5: CG iterations FP32c {res32.iterations}, FP64 {res64.iterations}; per column "
              f"{res32.column_iterations.tolist()}", flush=True)
        assert bool((res32.column_iterations > 0).all()) and bool((res64.column_iterations > 0).all())
        b = center(phen.y)
        bn = torch.linalg.norm(b, dim=0)
        h64 = shifted(op64, phen.delta)
        r32 = torch.linalg.norm(center(b - h64(res32.x)), dim=0)      # true residuals under the FP64 operator
        r64 = torch.linalg.norm(center(b - h64(res64.x)), dim=0)
        op_gap = torch.linalg.norm(shifted(op32, phen.delta)(res32.x) - h64(res32.x), dim=0)
        print(f"c5: true rel residual (FP64 operator) of FP32c solutions max {float((r32 / bn).max()):.2e}, of FP64 "
              f"solutions max {float((r64 / bn).max()):.2e}; operator gap at the FP32c solution "
              f"{float((op_gap / bn).max()):.2e}", flush=True)
        # FP64 sweeps: the true residual is CG's own (1e-6) up to FP64 rounding. FP32c sweeps: CG
        # converges on the FP32c operator, whose rounding (the operator gap, a few 1e-7 of ||b||, plus
        # its accumulation over the iterations' updates) sets the floor of the true FP64 residual.
        assert float((r64 / bn).max()) <= 2e-6
        assert float((r32 / bn).max()) <= 1e-5
        # H64 (x32 - x64) = r64 - r32 and H64 >= delta I on centred vectors: the solutions differ
        # by at most (||r32|| + ||r64||) / delta
        diff = torch.linalg.norm(res32.x - res64.x, dim=0)
        bound = (r32 + r64) / phen.delta
        print(f"c5: ||x32 - x64|| / bound max {float((diff / bound).max()):.3f}; relative distance max "
              f"{float((diff / torch.linalg.norm(res64.x, dim=0)).max()):.2e}", flush=True)
        assert bool((diff <= bound * 1.001).all())
