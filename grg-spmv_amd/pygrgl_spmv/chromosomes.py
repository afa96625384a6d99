"""Several GRGs (one per chromosome) in one HIP runtime, captured into one hipGraph.

    PYTHONPATH=grg-spmv_amd python -m pygrgl_spmv.chromosomes [--chromosomes 22 --samples 200000 --budget-gb 32 --ring 4]
    PYTHONPATH=grg-spmv_amd python -m pygrgl_spmv.chromosomes --streaming-budget

The batched configuration of the benchmark suite: per-chromosome synthetic GRGs
over the same samples with uneven mutation counts (U[20k, 400k]), one runtime
under a VRAM budget with a ring of streaming slots (``runtime.grgs`` as in the
reference's multi-GRG runtime, ``triton/backend.py:373-410``), mixed column
counts k in [1, 64] per GRG. Every GRG's up and down op (input copy, sweep,
output copy) is recorded into ONE hipGraph when every block is resident; the
graph is replayed and each replay checked bit for bit against the eager run.
The same sweeps then run from job tables (``HipRuntime.prepare_batched_sweep``):
level p of every GRG in one phase, as ONE persistent launch per direction with
grid-wide barriers between phases, and as one launch per phase; both are
checked against the per-GRG ops, captured and replayed. Under a budget that streams
blocks (``--streaming-budget``) the ops run eagerly and the batch runs as grouped
job tables with the streamed blocks copied and accumulated between phases.
(Parity of multi-GRG runtimes with the CPU oracle is covered by the tests.)

Prints progress on stderr and one JSON line on stdout.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def log(msg: str) -> None:
    print(f"[chromosomes] {msg}", file=sys.stderr, flush=True)


def mixed_k(i: int) -> int:
    """Column counts cycling through 1..64 (powers of two and odd sizes)."""
    return (1, 3, 8, 16, 5, 32, 64, 2, 12, 48, 7)[i % 11]


def streaming_budget(artifacts, *, ring: int = 4, fraction: float = 0.125) -> int:
    """A VRAM budget ``fraction`` of the way from the all-streamed layout's bytes to full
    residency: the planner keeps the largest blocks resident and streams the rest through
    the ring (the 32 GB of the configuration hold every block of these GRGs)."""
    from pygrgl_spmv.backends.hip import HipPlanPair, plan_hip_layout
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=64, max_k_down=64, need_down_miss_input=False, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=False)
    kw = dict(artifacts=[Path(a) for a in artifacts], pair=HipPlanPair.default(), dtype=np.float64, requirements=req,
              device=0, stream=0)
    full = plan_hip_layout(vram_budget_bytes=0, ring_buffer_size=0, **kw)
    streamed = plan_hip_layout(vram_budget_bytes=0, ring_buffer_size=int(ring), allow_residency=False, **kw)
    lo, hi = int(streamed.bytes_total), int(full.required_budget_for_full_residency)
    return lo + int((hi - lo) * float(fraction))


def run(artifacts, *, budget_bytes: int, ring: int, dtype=np.float64, replays: int = 3, seed: int = 4,
        autotune: bool = True, tune_grid: bool = True) -> dict:
    import torch

    from pygrgl_spmv.backends.hip import HipPlanPair, HipRuntime, plan_hip_layout
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=64, max_k_down=64, need_down_miss_input=False, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=False)
    with torch.cuda.device(0):
        cap = torch.cuda.Stream()
    t0 = time.perf_counter()
    layout = plan_hip_layout(artifacts=[Path(a) for a in artifacts], pair=HipPlanPair.default(), dtype=dtype, requirements=req,
                             vram_budget_bytes=int(budget_bytes), ring_buffer_size=int(ring), device=0, stream=cap)
    blocks = [b for a in layout.artifacts for b in (*a.blocks_up, *a.blocks_down)]
    out = {"grgs": len(artifacts), "budget_bytes": int(budget_bytes), "bytes_total": int(layout.bytes_total),
           "required_for_full_residency": int(layout.required_budget_for_full_residency),
           "blocks": len(blocks), "streamed_blocks": sum(not b.resident for b in blocks),
           "ring_slots": int(layout.allocated_ring_buffer_size)}
    log(f"layout: {out['bytes_total'] / 1e9:.2f} GB of {budget_bytes / 1e9:.1f} GB budget, "
        f"{out['streamed_blocks']} of {out['blocks']} blocks streamed")
    rng = np.random.default_rng(seed)
    with HipRuntime(layout, autotune=autotune) as rt:
        grgs = rt.grgs
        ops, xs, ys, res_up, res_dn = [], [], [], [], []
        edges = 0
        for i, g in enumerate(grgs):
            k = mixed_k(i)
            up = g.prepare_matmul_cuda(direction="up", k=k)
            dn = g.prepare_matmul_cuda(direction="down", k=k)
            ops.append((up, dn))
            xs.append(torch.from_numpy(rng.standard_normal((k, g.num_samples))).to(rt.device, up.input.dtype))
            ys.append(torch.from_numpy(rng.standard_normal((k, g.num_mutations))).to(rt.device, dn.input.dtype))
            res_up.append(torch.empty_like(up.output))
            res_dn.append(torch.empty_like(dn.output))
            edges += 2 * int(g.num_edges) * k
        out["ks"] = [mixed_k(i) for i in range(len(grgs))]
        out["edge_updates_per_pass"] = edges
        out["edges"] = [int(g.num_edges) for g in grgs]
        log(f"runtime entered and {2 * len(grgs)} ops prepared in {time.perf_counter() - t0:.1f}s")

        def all_ops():
            for (up, dn), x, y, ru, rd in zip(ops, xs, ys, res_up, res_dn):
                up.input.copy_(x)
                up()
                ru.copy_(up.output)
                dn.input.copy_(y)
                dn()
                rd.copy_(dn.output)

        with torch.cuda.stream(cap):
            all_ops()  # warm
        torch.cuda.synchronize()
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        with torch.cuda.stream(cap):
            ev[0].record(cap)
            all_ops()
            ev[1].record(cap)
        torch.cuda.synchronize()
        out["eager_ms"] = ev[0].elapsed_time(ev[1])
        out["eager_edge_updates_per_s"] = edges / (out["eager_ms"] / 1e3)
        eager_up = [r.cpu().clone() for r in res_up]
        eager_dn = [r.cpu().clone() for r in res_dn]
        capturable = all(up.capturable and dn.capturable for up, dn in ops)
        out["captured"] = bool(capturable)
        if capturable:
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=cap):
                all_ops()
            times = []
            for rep in range(int(replays)):
                for r in (*res_up, *res_dn):
                    r.zero_()
                torch.cuda.synchronize()
                with torch.cuda.stream(cap):
                    ev[0].record(cap)
                    graph.replay()
                    ev[1].record(cap)
                torch.cuda.synchronize()
                times.append(ev[0].elapsed_time(ev[1]))
                same = all(torch.equal(a.cpu(), b) for a, b in zip(res_up, eager_up)) and \
                    all(torch.equal(a.cpu(), b) for a, b in zip(res_dn, eager_dn))
                if not same:
                    raise AssertionError(f"graph replay {rep} differs from the eager run")
            out["graph_replay_ms"] = min(times)
            out["graph_replays_checked"] = int(replays)
            out["graph_edge_updates_per_s"] = edges / (min(times) / 1e3)
        log(f"eager {out['eager_ms']:.2f} ms" + (f", graph replay {out['graph_replay_ms']:.2f} ms" if capturable else
                                                  ", not captured (streamed blocks)"))
        if not capturable:
            # the same 44 sweeps as grouped job tables, the streamed blocks copied and
            # accumulated between phases (eager only)
            tol = 1e-12 if np.dtype(dtype) == np.float64 else 5e-7
            ups = rt.prepare_batched_sweep([dict(grg=g, direction="up", k=mixed_k(i)) for i, g in enumerate(grgs)],
                                           persistent=False)
            dns = rt.prepare_batched_sweep([dict(grg=g, direction="down", k=mixed_k(i)) for i, g in enumerate(grgs)],
                                           persistent=False)

            def batched_streamed():
                for v, x in zip(ups.inputs, xs):
                    v.copy_(x)
                for v, y in zip(dns.inputs, ys):
                    v.copy_(y)
                ups()
                dns()

            with torch.cuda.stream(cap):
                batched_streamed()
            torch.cuda.synchronize()
            err = 0.0
            for a, b in zip((*ups.outputs, *dns.outputs), (*eager_up, *eager_dn)):
                a64, b64 = a.cpu().double(), b.double()
                err = max(err, float((a64 - b64).norm() / max(float(b64.norm()), 1e-300)))
            if err > tol:
                raise AssertionError(f"grouped streamed batched sweep differs from the per-GRG ops: {err:.2e}")
            with torch.cuda.stream(cap):
                ev[0].record(cap)
                batched_streamed()
                ev[1].record(cap)
            torch.cuda.synchronize()
            res = {"phases_up": ups.num_phases, "phases_down": dns.num_phases, "jobs": ups.num_jobs + dns.num_jobs,
                   "streamed_block_copies": len(ups._streamed) + len(dns._streamed),
                   "eager_ms": ev[0].elapsed_time(ev[1]), "max_rel_err_vs_per_grg_ops": err}
            res["eager_edge_updates_per_s"] = edges / (res["eager_ms"] / 1e3)
            out["batched_grouped_streamed"] = res
            log(f"batched (grouped, streamed): {res['jobs']} jobs, {res['streamed_block_copies']} block copies, eager "
                f"{res['eager_ms']:.2f} ms, rel err vs per-GRG ops {err:.1e}")
            del ups, dns
        if capturable:
            # the same 44 sweeps from job tables: all GRGs' level p in one phase, the up and
            # down sweeps as two batches, or as ONE batch (every sweep writes private
            # buffers, so an up and a down sweep share phases and barriers)
            tol = 1e-12 if np.dtype(dtype) == np.float64 else 5e-7
            ins = [*xs, *ys]
            want = (*eager_up, *eager_dn)
            up_entries = [dict(grg=g, direction="up", k=mixed_k(i)) for i, g in enumerate(grgs)]
            dn_entries = [dict(grg=g, direction="down", k=mixed_k(i)) for i, g in enumerate(grgs)]
            for mode, persistent, combined in (("persistent", True, False), ("grouped", False, False),
                                               ("persistent_one_batch", True, True), ("grouped_one_batch", False, True)):
                if combined:
                    batches = [rt.prepare_batched_sweep(up_entries + dn_entries, persistent=persistent)]
                else:
                    batches = [rt.prepare_batched_sweep(up_entries, persistent=persistent),
                               rt.prepare_batched_sweep(dn_entries, persistent=persistent)]
                views = [v for b in batches for v in b.inputs]
                outputs = [o for b in batches for o in b.outputs]
                grid_ms = {}
                if persistent and tune_grid:
                    grid_ms = [b.tune_grid() for b in batches]
                    log(f"{mode}: persistent grids (ms per batch) {grid_ms} -> {[b.grid for b in batches]}")

                def batched():
                    for v, x in zip(views, ins):
                        v.copy_(x)
                    for b in batches:
                        b()

                with torch.cuda.stream(cap):
                    batched()
                torch.cuda.synchronize()
                for b in batches:
                    b.check()
                err = 0.0
                for a, b in zip(outputs, want):
                    a64 = a.cpu().double()
                    b64 = b.double()
                    err = max(err, float((a64 - b64).norm() / max(float(b64.norm()), 1e-300)))
                if err > tol:
                    raise AssertionError(f"{mode} batched sweep differs from the per-GRG ops: {err:.2e}")
                with torch.cuda.stream(cap):
                    ev[0].record(cap)
                    batched()
                    ev[1].record(cap)
                torch.cuda.synchronize()
                res = {"batches": len(batches), "phases": [b.num_phases for b in batches],
                       "jobs": sum(b.num_jobs for b in batches),
                       "grid": [b.grid for b in batches] if persistent else None, "grid_ms": grid_ms,
                       "eager_ms": ev[0].elapsed_time(ev[1]), "max_rel_err_vs_per_grg_ops": err}
                eager_b = [o.cpu().clone() for o in outputs]
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph, stream=cap):
                    batched()
                times = []
                for rep in range(int(replays)):
                    for o in outputs:
                        o.zero_()
                    torch.cuda.synchronize()
                    with torch.cuda.stream(cap):
                        ev[0].record(cap)
                        graph.replay()
                        ev[1].record(cap)
                    torch.cuda.synchronize()
                    times.append(ev[0].elapsed_time(ev[1]))
                    for b in batches:
                        b.check()
                    if not all(torch.equal(o.cpu(), e) for o, e in zip(outputs, eager_b)):
                        raise AssertionError(f"{mode} batched graph replay {rep} differs from its eager run")
                res["graph_replay_ms"] = min(times)
                res["graph_edge_updates_per_s"] = edges / (min(times) / 1e3)
                if persistent:
                    res.update(batches[0].barrier_costs())
                    log(f"{mode}: persistent kernel at {batches[0].grid} workgroups: grid barrier {res['barrier_us']:.2f} us, "
                        f"one empty phase (reset, launch, registration) {res['one_phase_us']:.2f} us")
                out[f"batched_{mode}"] = res
                log(f"batched ({mode}): {res['jobs']} jobs in {res['phases']} phases, eager {res['eager_ms']:.2f} ms, "
                    f"graph replay {res['graph_replay_ms']:.2f} ms, rel err vs per-GRG ops {err:.1e}")
                del graph, batches
    out["seconds_total"] = time.perf_counter() - t0
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--chromosomes", type=int, default=22)
    ap.add_argument("--samples", type=int, default=200_000)
    ap.add_argument("--budget-gb", type=float, default=32.0)
    ap.add_argument("--streaming-budget", action="store_true",
                    help="instead of --budget-gb: a budget an eighth of the way from all-streamed to full residency")
    ap.add_argument("--ring", type=int, default=4)
    ap.add_argument("--artifact-dir", default="/tmp/grg_chrom")
    ap.add_argument("--fp32c", action="store_true")
    args = ap.parse_args()

    from pygrgl_spmv.synth import chromosome_workloads, write_artifact

    d = Path(args.artifact_dir)
    d.mkdir(parents=True, exist_ok=True)
    arts = []
    t = time.perf_counter()
    for i, cfg in enumerate(chromosome_workloads(args.chromosomes, args.samples)):
        p = d / f"chr{i + 1}-{args.samples}-seed{cfg.seed}.grg_spmv"
        if not p.exists():
            write_artifact(cfg, p)
        arts.append(p)
    log(f"{len(arts)} artifacts ready in {time.perf_counter() - t:.1f}s")
    budget = streaming_budget(arts, ring=args.ring) if args.streaming_budget else int(args.budget_gb * 1e9)
    res = run(arts, budget_bytes=budget, ring=args.ring, dtype=np.float32 if args.fp32c else np.float64)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
