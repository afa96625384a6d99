"""GRG matmul benchmark: edge-updates/s of one up + down sweep pair on one MI355X.

A step is one up sweep (samples -> mutations) followed by one down sweep
(mutations -> samples, with the missing-data input when the workload has
missingness) over one batch of k synthetic N(0,1) columns, through the drop-in
op surface (``grg.prepare_matmul_cuda``), with inputs resident in HBM. Each
edge-update is one (edge, column) add, so a step performs 2 * E * k.

The default workload is the 500,000-sample x 5,000,000-mutation synthetic GRG
(5 % missing genotypes, coalescence counts kept, k = 16), in FP64 and in the
FP32-compensated form; it is generated in-process from a fixed seed and
written as a ``.grg_spmv`` artifact (no datasets are downloaded).

Timing: HIP events recorded on the stream the sweep kernels are launched on
(``HipRuntime.sweep_timer``), bracketing all launches of one sweep; the
roofline figure divides the sweep's algorithmic bytes by that time.

Prints ONE JSON line. ``--gpus`` above 1 is rejected: this build runs on one
GPU.
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
METRIC = "GRG matmul edge-updates/s (k=16, up+down), % HBM roofline; FP32c vs FP64"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(stats: dict, k: int, w: int, direction: str, use_miss: bool) -> int:
    """Compulsory HBM traffic of one sweep: the structure once (int32 columns and
    row pointers of the merged per-level CSR), every node value written once and
    read once, the endpoint input read and output written once, the selector once
    (and the missingness selector plus its M x k input when the down sweep takes
    one)."""
    E, N, S, M, L = stats["edges"], stats["nodes"], stats["samples"], stats["mutations"], stats["levels"]
    R = N + L
    n_in, n_out = (S, M) if direction == "up" else (M, S)
    b = 4 * E + 4 * R + 2 * N * k * w + (n_in + n_out) * k * w + 4 * (stats["sel_mut_nnz"] + M + 1)
    if use_miss:
        b += 4 * (stats["sel_miss_nnz"] + M + 1) + M * k * w
    return int(b)


def cpu_threads() -> int:
    """Threads for the CPU baseline: the cores this process may use, capped by OMP_NUM_THREADS."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def host_cores() -> dict:
    """Core counts of this host: logical (``os.cpu_count()``, what the reference's MklRuntime uses
    for ``n_threads=0``, ``mkl/backend.py:296-300``), physical (psutil), and the cores this
    process may run on (its affinity mask)."""
    try:
        import psutil

        physical = psutil.cpu_count(logical=False)
    except Exception:  # psutil absent or the count unknown
        physical = None
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"logical_cores": os.cpu_count(), "physical_cores": physical, "affinity_cores": affinity}


def run_gpu(artifact: Path, *, k: int, dtype, steps: int, warmup: int, use_miss: bool, calibrate_mib: int = 0,
            runtime_kwargs: dict | None = None, graph: bool = False):
    import torch

    from pygrgl_spmv.backends.hip import HipPlanPair, HipRuntime, plan_hip_layout
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=use_miss, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=use_miss)
    # the runtime's caller stream (the stream its ops join and a capture records on)
    caller = torch.cuda.Stream(device=torch.device("cuda", 0)) if graph else 0
    layout = plan_hip_layout(artifacts=[artifact], pair=HipPlanPair.default(), dtype=dtype, requirements=req,
                             vram_budget_bytes=0, ring_buffer_size=0, device=0, stream=caller)
    t0 = time.perf_counter()
    with HipRuntime(layout, **(runtime_kwargs or {})) as rt:
        (grg,) = rt.grgs
        up = grg.prepare_matmul_cuda(direction="up", k=k)
        down = grg.prepare_matmul_cuda(direction="down", k=k, use_miss=use_miss)
        log(f"{np.dtype(dtype).name}: runtime entered and ops prepared in {time.perf_counter() - t0:.1f}s")
        rng = np.random.default_rng(2026)
        dev = up.input.device
        x = torch.from_numpy(rng.standard_normal((k, grg.num_samples)).astype(dtype)).to(dev)
        y = torch.from_numpy(rng.standard_normal((k, grg.num_mutations)).astype(dtype)).to(dev)
        miss = torch.from_numpy(rng.standard_normal((k, grg.num_mutations)).astype(dtype)).to(dev) if use_miss else None

        def load_inputs(which):
            # up's input and down's output share the arena (and down's input is up's output):
            # inputs are restored outside the timed events, on the runtime's caller stream
            with torch.cuda.stream(rt.stream):
                if which == "up":
                    up.input.copy_(x)
                else:
                    down.input.copy_(y)
                    if miss is not None:
                        down.miss_input.copy_(miss)

        for _ in range(warmup):
            load_inputs("up")
            up()
            load_inputs("down")
            down()
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        with rt.sweep_timer() as ev:
            for _ in range(steps):
                load_inputs("up")
                up()
                load_inputs("down")
                down()
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        t_up = sum(a.elapsed_time(b) for a, b in ev[0::2]) / steps / 1e3
        t_dn = sum(a.elapsed_time(b) for a, b in ev[1::2]) / steps / 1e3
        t_graph = None
        if graph:
            # the same up + down pair captured once into a hipGraph and replayed: one
            # launch of the whole step (the values are not checked here; the clean pass
            # below runs eagerly)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=caller):
                up()
                down()
            for _ in range(max(1, warmup)):
                g.replay()
            torch.cuda.synchronize()
            gev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
            with torch.cuda.stream(caller):
                gev[0].record(caller)
                for i in range(steps):
                    g.replay()
                    gev[i + 1].record(caller)
            torch.cuda.synchronize()
            t_graph = min(gev[i].elapsed_time(gev[i + 1]) for i in range(steps)) / 1e3
            t_graph_mean = gev[0].elapsed_time(gev[-1]) / steps / 1e3
            log(f"{np.dtype(dtype).name}: captured step replay {t_graph * 1e3:.3f} ms (best), {t_graph_mean * 1e3:.3f} ms (mean)")
            del g
        # one clean pass for the accuracy report
        load_inputs("up")
        up()
        torch.cuda.synchronize()
        up_out = up.output.detach().cpu().numpy().astype(np.float64)
        load_inputs("down")
        down()
        torch.cuda.synchronize()
        down_out = down.output.detach().cpu().numpy().astype(np.float64)
        if calibrate_mib > 0:
            # a plain device copy of known size, for scaling HBM byte counters
            n = calibrate_mib * (1 << 20) // 8
            a = torch.empty(n, dtype=torch.float64, device=dev).fill_(1.0)
            b = torch.empty_like(a)
            torch.cuda.synchronize()
            b.copy_(a)
            torch.cuda.synchronize()
            log(f"calibration copy: {calibrate_mib} MiB read + {calibrate_mib} MiB written")
            del a, b
        tune = [{"config": key[1], "level": key[2], "k": key[3], **{n: (list(v) if isinstance(v, tuple) else v) for n, v in rec.items()}}
                for key, rec in rt.tune_log.items()]
    return {"t_up": t_up, "t_down": t_dn, "wall_per_step": wall / steps, "up_out": up_out, "down_out": down_out,
            "tune": tune, "t_graph": t_graph}


def cpu_baseline(artifact: Path, stats: dict, *, k: int, use_miss: bool, budget_s: float = 20.0):
    """The reference's MklRuntime sweep loop (restated in oracle/mkl_baseline.py) on this host's cores."""
    from oracle import grg_oracle as O
    from oracle.mkl_baseline import MklSweep

    g = O.load_artifact(artifact)
    sweep = MklSweep(g, max_k=k, n_threads=cpu_threads())
    try:
        rng = np.random.default_rng(2026)
        x = rng.standard_normal((g.num_samples, k))
        y = rng.standard_normal((g.num_mutations, k))
        miss = rng.standard_normal((g.num_mutations, k)) if use_miss else None
        v_up = np.zeros((g.num_nodes, k))
        v_dn = np.zeros((g.num_nodes, k))
        seed_dn = g.sel_mut.T @ y
        if miss is not None:
            seed_dn = seed_dn + g.sel_miss.T @ miss

        def one():
            v_up[:] = 0.0
            v_up[: g.num_samples] = x
            t0 = time.perf_counter()
            sweep.sweep(v_up, "up")
            t1 = time.perf_counter()
            v_dn[:] = seed_dn
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
        "sample": f"MKL per-block level sweep (mkl_sparse_d_mm, up N/CSR + down T on the same handles, k={k}) over the "
                  f"whole workload, mean of {len(times)} up+down steps after 1 warmup; seeding and selector gathers "
                  f"outside the timed sweep; {sweep.n_threads} threads; library {sweep.path}",
        "ms_per_step": per_step * 1e3,
        **host_cores(),
    }


def gather_ceiling(k: int = 16, rows: int = 4_000_000, degree: int = 8, reps: int = 5) -> list:
    """Throughput of the sweep kernel (``row_gather_kernel``) on synthetic single-level
    launches: ``rows`` output rows of ``degree`` entries each gather FP64 k-wide source
    rows, from a source array of 64 MB or 3.3 GB, with indices either uniformly random
    or confined to a 65,536-row window that moves with the output row. No GRG structure:
    this is the gather rate the hardware gives this kernel, the ceiling the sweep levels
    are compared with."""
    import torch

    from pygrgl_spmv.backends.hip import ffi

    lib = ffi.hip_library()
    dev = torch.device("cuda", 0)
    out_rows = []
    ptr = torch.arange(0, (rows + 1) * degree, degree, dtype=torch.int32, device=dev)
    out = torch.empty((rows, k), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(2026)
    for src_rows in (500_000, 26_000_000):
        src = torch.ones((src_rows, k), dtype=torch.float64, device=dev)
        for pattern in ("random", "window"):
            if pattern == "random":
                idx = torch.randint(0, src_rows, (rows * degree,), dtype=torch.int64, device=dev, generator=gen)
            else:
                base = torch.arange(rows * degree, device=dev, dtype=torch.int64) // degree * src_rows // rows
                idx = (base + torch.randint(0, 65536, (rows * degree,), device=dev, generator=gen)) % src_rows
            idx = idx.to(torch.int32)
            for unroll, es in ((8, 1), (16, 1)):
                a = ffi.GrgHipGatherArgs()
                a.dtype, a.accumulate, a.row_begin, a.row_end, a.k = ffi.F64, 0, 0, rows, k
                a.cpl, a.kl, a.es, a.unroll, a.rows_per_group, a.flags = 2, k // 2, es, unroll, 1, 0
                a.out, a.ld_out, a.out_base = out.data_ptr(), k, 0
                a.nseg, a.nmap = 1, 0
                a.seg[0].ptr, a.seg[0].idx, a.seg[0].src, a.seg[0].ld, a.seg[0].ptr_base = ptr.data_ptr(), idx.data_ptr(), src.data_ptr(), k, 0
                a.heavy_threshold, a.heavy_rows, a.num_heavy = -1, 0, 0
                lib.gather(a, stream.cuda_stream)
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(reps + 1)]
                ev[0].record(stream)
                for i in range(reps):
                    lib.gather(a, stream.cuda_stream)
                    ev[i + 1].record(stream)
                ev[-1].synchronize()
                ms = min(ev[i].elapsed_time(ev[i + 1]) for i in range(reps))
                gathers = rows * degree
                rec = {"source_MB": src_rows * k * 8 / 1e6, "pattern": pattern, "unroll": unroll, "ms": ms,
                       "G_gathers_per_s": gathers / ms / 1e6, "gathered_TB_per_s": gathers * k * 8 / ms / 1e9,
                       "written_TB_per_s": rows * k * 8 / ms / 1e9}
                log(f"gather ceiling: {rec}")
                out_rows.append(rec)
            del idx
        del src
    return out_rows


def load_traffic(workload: str, k: int):
    """HBM bytes per sweep measured by rocprofv3 PMC runs of this benchmark (FETCH_SIZE and
    WRITE_SIZE over the sweep's dispatches, scaled by a plain copy), if recorded."""
    p = ROOT / "profiles" / f"row_gather_{workload}_k{k}_traffic.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return None


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--artifact-dir", default=os.environ.get("GRG_BENCH_DIR", "/tmp/grg_bench"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32c", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="FP32c only (profiling runs)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed region, run one device copy of this many MiB (for PMC scaling)")
    ap.add_argument("--tune-log", default="",
                    help="write the per-level tuner choices and launch times of each form to this JSON file")
    ap.add_argument("--runtime-option", action="append", default=[], metavar="NAME=VALUE",
                    help="keyword argument of HipRuntime for both value types, e.g. panels=True (repeatable)")
    ap.add_argument("--graph", action="store_true",
                    help="also time the up + down pair captured into one hipGraph and replayed")
    ap.add_argument("--gather-ceiling", action="store_true",
                    help="instead of the sweep: the sweep kernel's gather rate on synthetic random and windowed indices")
    args = ap.parse_args()
    if args.gather_ceiling:
        print(json.dumps({"gather_ceiling": gather_ceiling(k=int(args.k))}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        return 2

    from pygrgl_spmv.grg.artifact import scan_grg_spmv
    from pygrgl_spmv.synth import WORKLOADS, write_artifact

    cfg = WORKLOADS[args.workload]
    art_dir = Path(args.artifact_dir)
    art_dir.mkdir(parents=True, exist_ok=True)
    artifact = art_dir / f"{args.workload}-seed{cfg.seed}.grg_spmv"
    t0 = time.perf_counter()
    if not artifact.exists():
        log(f"generating workload {args.workload} into {artifact}")
        write_artifact(cfg, artifact)
    scan = scan_grg_spmv(artifact)
    stats = {"samples": scan.num_samples, "mutations": scan.num_mutations, "nodes": scan.num_nodes,
             "edges": scan.num_edges, "levels": scan.num_levels, "sel_mut_nnz": scan.selector_mut_nnz,
             "sel_miss_nnz": scan.selector_miss_nnz}
    use_miss = bool(scan.has_missing_data) and stats["sel_miss_nnz"] > 0
    log(f"workload {args.workload}: {stats} ready in {time.perf_counter() - t0:.1f}s; down sweep uses miss input: {use_miss}")

    import ast

    runtime_kwargs = {}
    for opt in args.runtime_option:
        name, _, value = opt.partition("=")
        runtime_kwargs[name.strip()] = ast.literal_eval(value.strip())
    if runtime_kwargs:
        log(f"runtime options: {runtime_kwargs}")
    k = int(args.k)
    eu = 2 * stats["edges"] * k
    variants = {}
    f64 = None
    if not args.no_fp64:
        f64 = run_gpu(artifact, k=k, dtype=np.float64, steps=args.steps, warmup=args.warmup, use_miss=use_miss,
                      calibrate_mib=args.calibrate_copy_mib, runtime_kwargs=runtime_kwargs, graph=args.graph)
        t64 = f64["t_up"] + f64["t_down"]
        b64 = algorithmic_bytes(stats, k, 8, "up", False) + algorithmic_bytes(stats, k, 8, "down", use_miss)
        log(f"fp64: up {f64['t_up']*1e3:.3f} ms, down {f64['t_down']*1e3:.3f} ms, {eu / t64 / 1e9:.2f} G edge-updates/s, "
            f"{b64 / t64 / 1e9:.0f} GB/s algorithmic")
        variants["fp64"] = {"value": eu / t64, "ms_per_step": t64 * 1e3, "ms_up": f64["t_up"] * 1e3,
                            "ms_down": f64["t_down"] * 1e3, "algorithmic_bytes": b64, "algorithmic_GBs": b64 / t64 / 1e9,
                            "wall_ms_per_step": f64["wall_per_step"] * 1e3}
        if f64["t_graph"] is not None:
            variants["fp64"]["graph_ms_per_step"] = f64["t_graph"] * 1e3
    if not args.no_fp32c:
        f32 = run_gpu(artifact, k=k, dtype=np.float32, steps=args.steps, warmup=args.warmup, use_miss=use_miss,
                      calibrate_mib=args.calibrate_copy_mib if args.no_fp64 else 0, runtime_kwargs=runtime_kwargs,
                      graph=args.graph)
        t32 = f32["t_up"] + f32["t_down"]
        b32 = algorithmic_bytes(stats, k, 4, "up", False) + algorithmic_bytes(stats, k, 4, "down", use_miss)
        v = {"value": eu / t32, "ms_per_step": t32 * 1e3, "ms_up": f32["t_up"] * 1e3, "ms_down": f32["t_down"] * 1e3,
             "algorithmic_bytes": b32, "algorithmic_GBs": b32 / t32 / 1e9, "wall_ms_per_step": f32["wall_per_step"] * 1e3}
        if f32["t_graph"] is not None:
            v["graph_ms_per_step"] = f32["t_graph"] * 1e3
        msg = f"fp32c: up {f32['t_up']*1e3:.3f} ms, down {f32['t_down']*1e3:.3f} ms, {eu / t32 / 1e9:.2f} G edge-updates/s"
        if f64 is not None:
            def col_err(a, b):
                # per-column relative error (max-norm over rows) against the FP64 result
                return np.abs(a - b).max(axis=1) / np.maximum(np.abs(b).max(axis=1), 1e-300)

            e_up = col_err(f32["up_out"], f64["up_out"])
            e_dn = col_err(f32["down_out"], f64["down_out"])
            v.update({"speedup_vs_fp64": (f64["t_up"] + f64["t_down"]) / t32,
                      "max_rel_err_vs_fp64": float(max(e_up.max(), e_dn.max())),
                      "per_column_rel_err_up": [float(e) for e in e_up],
                      "per_column_rel_err_down": [float(e) for e in e_dn]})
            msg += f", {v['speedup_vs_fp64']:.2f}x fp64, max per-column rel err {v['max_rel_err_vs_fp64']:.2e}"
        log(msg)
        variants["fp32c"] = v
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(artifact, stats, k=k, use_miss=use_miss)
            log(f"cpu baseline: {cpu['value'] / 1e9:.3f} G edge-updates/s on {cpu['cores']} threads")
        except Exception as exc:  # reported, not fatal
            cpu = {"value": None, "unit": "edge-updates/s", "cores": cpu_threads(), "kind": "port", "sample": f"unavailable: {exc}"}
    if args.tune_log:
        with open(args.tune_log, "w") as fh:
            json.dump({"fp64": f64["tune"] if f64 is not None else None,
                       "fp32c": f32["tune"] if not args.no_fp32c else None}, fh, indent=1)
    head = variants.get("fp64") or variants["fp32c"]
    traffic = load_traffic(args.workload, k)
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "edge-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if "fp64" in variants else "f32c",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: synthetic GRG {stats['samples']} samples x {stats['mutations']} mutations "
                               f"({stats['nodes']} nodes, {stats['edges']} edges, {stats['levels']} levels), k={k}, up + down"
                               + (" with missing-data input" if use_miss else ""),
                   "k": k, **stats},
        "roofline": {"bound": "hbm", "achieved": head["algorithmic_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["algorithmic_GBs"] / HBM_PEAK_GBS,
                     "traffic": None if traffic is None else traffic.get("bytes_per_step")},
        "variants": variants,
        "cpu_baseline": cpu,
    }
    if traffic is not None:
        out["roofline"]["traffic_source"] = traffic.get("source")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
