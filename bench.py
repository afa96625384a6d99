"""GRG matmul benchmark: edge-updates/s of one up + down sweep pair on one MI355X.

A step is one up sweep (samples -> mutations) followed by one down sweep
(mutations -> samples) over one batch of k synthetic N(0,1) columns, through
the drop-in op surface (``grg.prepare_matmul_cuda``), with inputs resident in
HBM. Each edge-update is one (edge, column) add, so a step performs 2 * E * k.
The workload is a seeded synthetic GRG generated in-process and written as a
``.grg_spmv`` artifact (no datasets are downloaded).

Prints ONE JSON line (rank 0). ``--gpus`` above 1 is rejected: this build runs
on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "grg-spmv_amd", ROOT):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats: dict, k: int, w: int, direction: str) -> int:
    """Compulsory traffic of one sweep: structure once, each node value written and read once,
    endpoints read/written once, selector once."""
    E, N, S, M, L, sel = stats["edges"], stats["nodes"], stats["samples"], stats["mutations"], stats["levels"], stats["sel_mut_nnz"]
    R = N + L
    n_in, n_out = (S, M) if direction == "up" else (M, S)
    return int(4 * E + 4 * R + 2 * N * k * w + (n_in + n_out) * k * w + 4 * (sel + M + 1))


def run_gpu(artifact: Path, stats: dict, *, k: int, dtype, steps: int, warmup: int):
    import torch

    from pygrgl_spmv.backends.hip import HipPlanPair, HipRuntime, plan_hip_layout
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=False, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=False)
    layout = plan_hip_layout(artifacts=[artifact], pair=HipPlanPair.default(), dtype=dtype, requirements=req,
                             vram_budget_bytes=0, ring_buffer_size=0, device=0, stream=0)
    with HipRuntime(layout) as rt:
        (grg,) = rt.grgs
        up = grg.prepare_matmul_cuda(direction="up", k=k)
        down = grg.prepare_matmul_cuda(direction="down", k=k)
        rng = np.random.default_rng(2026)
        x = torch.from_numpy(rng.standard_normal((k, grg.num_samples)).astype(dtype)).to(up.input.device)
        up.input.copy_(x)
        caller = rt.stream
        for _ in range(warmup):
            up()
            down()
            up.input.copy_(x)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        for a, b, c in ev:
            # down's output aliases up's input (shared arena); restore the batch outside the timed events
            up.input.copy_(x)
            a.record(caller)
            up()
            b.record(caller)
            down()
            c.record(caller)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        t_up = sum(a.elapsed_time(b) for a, b, _ in ev) / steps / 1e3
        t_dn = sum(b.elapsed_time(c) for _, b, c in ev) / steps / 1e3
        # one clean pass for the accuracy report
        up.input.copy_(x)
        up()
        up_out = up.output.detach().cpu().numpy().astype(np.float64)
        down()
        down_out = down.output.detach().cpu().numpy().astype(np.float64)
    return {"t_up": t_up, "t_down": t_dn, "wall_per_step": wall / steps, "up_out": up_out, "down_out": down_out}


def cpu_baseline(artifact: Path, stats: dict, *, k: int, budget_s: float = 20.0):
    """The reference's MklRuntime sweep loop (restated in oracle/mkl_baseline.py) on this host's cores."""
    from oracle import grg_oracle as O
    from oracle.mkl_baseline import MklSweep

    g = O.load_artifact(artifact)
    sweep = MklSweep(g, max_k=k)
    try:
        rng = np.random.default_rng(2026)
        x = rng.standard_normal((g.num_samples, k))
        y = rng.standard_normal((g.num_mutations, k))
        v_up = np.zeros((g.num_nodes, k))
        v_dn = np.zeros((g.num_nodes, k))

        def one():
            v_up[:] = 0.0
            v_up[: g.num_samples] = x
            t0 = time.perf_counter()
            sweep.sweep(v_up, "up")
            t1 = time.perf_counter()
            v_dn[:] = 0.0
            v_dn[:] = g.sel_mut.T @ y
            t2 = time.perf_counter()
            sweep.sweep(v_dn, "down")
            t3 = time.perf_counter()
            return (t1 - t0) + (t3 - t2)

        one()  # warmup
        times = []
        t_start = time.perf_counter()
        while len(times) < 10 and (time.perf_counter() - t_start) < budget_s:
            times.append(one())
        per_step = float(np.mean(times))
    finally:
        sweep.close()
    return {
        "value": 2 * stats["edges"] * k / per_step,
        "unit": "edge-updates/s",
        "cores": int(sweep.n_threads),
        "kind": "port",
        "sample": f"MKL per-block sweep (up + down, k={k}, mean of {len(times)} steps after 1 warmup) on the full workload; library {sweep.path}",
        "ms_per_step": per_step * 1e3,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--artifact-dir", default=os.environ.get("GRG_BENCH_DIR", "/tmp/grg_bench"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32c", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        return 2

    from pygrgl_spmv.synth import WORKLOADS, write_artifact

    cfg = WORKLOADS[args.workload]
    art_dir = Path(args.artifact_dir)
    art_dir.mkdir(parents=True, exist_ok=True)
    artifact = art_dir / f"{args.workload}-seed{cfg.seed}.grg_spmv"
    t0 = time.perf_counter()
    if not artifact.exists():
        write_artifact(cfg, artifact)
    from pygrgl_spmv.grg.artifact import scan_grg_spmv

    scan = scan_grg_spmv(artifact)
    stats = {"samples": scan.num_samples, "mutations": scan.num_mutations, "nodes": scan.num_nodes,
             "edges": scan.num_edges, "levels": scan.num_levels, "sel_mut_nnz": scan.selector_mut_nnz}
    log(f"workload {args.workload}: {stats} ready in {time.perf_counter() - t0:.1f}s")

    k = int(args.k)
    f64 = run_gpu(artifact, stats, k=k, dtype=np.float64, steps=args.steps, warmup=args.warmup)
    eu = 2 * stats["edges"] * k
    t64 = f64["t_up"] + f64["t_down"]
    b64 = algorithmic_bytes(stats, k, 8, "up") + algorithmic_bytes(stats, k, 8, "down")
    log(f"fp64: up {f64['t_up']*1e3:.3f} ms, down {f64['t_down']*1e3:.3f} ms, {eu / t64 / 1e9:.2f} G edge-updates/s, {b64 / t64 / 1e9:.0f} GB/s")
    variants = {"fp64": {"value": eu / t64, "ms_per_step": t64 * 1e3, "ms_up": f64["t_up"] * 1e3, "ms_down": f64["t_down"] * 1e3,
                         "algorithmic_GBs": b64 / t64 / 1e9}}
    if not args.no_fp32c:
        f32 = run_gpu(artifact, stats, k=k, dtype=np.float32, steps=args.steps, warmup=args.warmup)
        t32 = f32["t_up"] + f32["t_down"]
        b32 = algorithmic_bytes(stats, k, 4, "up") + algorithmic_bytes(stats, k, 4, "down")

        def rel(a, b):
            return float((np.abs(a - b).max(axis=1) / np.maximum(np.abs(b).max(axis=1), 1e-300)).max())

        err = max(rel(f32["up_out"], f64["up_out"]), rel(f32["down_out"], f64["down_out"]))
        log(f"fp32c: {eu / t32 / 1e9:.2f} G edge-updates/s, {b32 / t32 / 1e9:.0f} GB/s, rel err vs fp64 {err:.2e}")
        variants["fp32c"] = {"value": eu / t32, "ms_per_step": t32 * 1e3, "ms_up": f32["t_up"] * 1e3, "ms_down": f32["t_down"] * 1e3,
                             "algorithmic_GBs": b32 / t32 / 1e9, "speedup_vs_fp64": t64 / t32, "max_rel_err_vs_fp64": err}
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(artifact, stats, k=k)
            log(f"cpu baseline: {cpu['value'] / 1e9:.3f} G edge-updates/s on {cpu['cores']} threads")
        except Exception as exc:  # reported, not fatal
            cpu = {"value": None, "unit": "edge-updates/s", "cores": os.cpu_count(), "kind": "port", "sample": f"unavailable: {exc}"}
    out = {
        "metric": "GRG matmul edge-updates/s (k=16, up+down), % HBM roofline; FP32c vs FP64",
        "value": variants["fp64"]["value"],
        "unit": "edge-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": variants["fp64"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: synthetic GRG {stats['samples']} samples x {stats['mutations']} mutations "
                               f"({stats['nodes']} nodes, {stats['edges']} edges, {stats['levels']} levels), k={k}, up+down",
                   "k": k, **stats},
        "roofline": {"bound": "hbm", "achieved": b64 / t64 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": b64 / t64 / 1e9 / HBM_PEAK_GBS, "traffic": None},
        "variants": variants,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
