"""Mixed-model CG with the learned warm start on a synthetic GRG, end to end on one GPU.

    python -m pygrgl_spmv.lmm.experiment --workload c5 [--n-train 2048 --n-test 256 --k 64]

Steps (all on the GPU except the artifact generation):

1. the standardized kinship operator K over the workload's GRG, sweeps in the
   FP32-compensated form (or FP64 with ``--fp64``);
2. the top-r eigenpairs of K (block subspace iteration, FP64 re-orthogonalisation);
3. ``n_train`` synthetic phenotypes solved by plain zero-start CG to the
   tolerance: the network's training set;
4. the warm-start network trained on them (HIP forward / backward / Adam);
5. ``n_test`` fresh phenotypes solved to the same residual (||r|| <= tol ||P y||)
   from zero, from the closed-form eigenvalue start and from the network's
   start; iterations per phenotype are compared. A fourth solve starts from the
   best coefficients of the same form (fitted to the zero-start solution): the
   bound any predictor of those coefficients can reach on this data.

Prints progress lines on stderr and one JSON line on stdout.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np


def log(msg: str) -> None:
    print(f"[lmm] {msg}", file=sys.stderr, flush=True)


def solve_batches(op, phen, *, rel_tol, max_iter, start=None):
    """Solve every k-column batch of ``phen``; ``start(y, delta, first_column)`` gives x0 (None: zero start)."""
    import torch

    from pygrgl_spmv.lmm.cg import conj_grad, shifted

    xs, cols, apps, rels, batch_its = [], [], 0, [], []
    for s in range(0, phen.y.shape[1], op.k):
        y = phen.y[:, s:s + op.k]
        d = phen.delta[s:s + op.k]
        x0 = None if start is None else start(y, d, s)
        res = conj_grad(shifted(op, d), y, x0=x0, rel_tol=rel_tol, max_iter=max_iter)
        xs.append(res.x)
        cols.append(res.column_iterations)
        rels.append(res.rel_residual)
        apps += res.applications
        batch_its.append(res.iterations)
    return torch.cat(xs, dim=1), torch.cat(cols).cpu().numpy(), torch.cat(rels).cpu().numpy(), apps, batch_its


def run(artifacts, *, k: int = 64, r: int = 20, n_train: int = 2048, n_test: int = 256, rel_tol: float = 1e-6,
        max_iter: int = 1000, dtype=np.float32, epochs: int = 300, seed: int = 2026) -> dict:
    import torch

    from pygrgl_spmv.backends.hip import HipPlanPair, HipRuntime, plan_hip_layout
    from pygrgl_spmv.grg import RuntimeRequirements
    from pygrgl_spmv.lmm.operator import StandardizedGrgOperator
    from pygrgl_spmv.lmm.pcs import top_eigenpairs
    from pygrgl_spmv.lmm.simulate import Phenotypes, simulate
    from pygrgl_spmv.lmm.warmstart import Basis, WarmStartNet, closed_form_coefficients, start_from_coefficients, targets

    req = RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=False, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=True)
    layout = plan_hip_layout(artifacts=[Path(a) for a in artifacts], pair=HipPlanPair.default(), dtype=dtype,
                             requirements=req, vram_budget_bytes=0, ring_buffer_size=0, device=0, stream=0)
    out = {"k": k, "r": r, "n_train": n_train, "n_test": n_test, "rel_tol": rel_tol, "sweep_dtype": np.dtype(dtype).name}
    t0 = time.perf_counter()
    with HipRuntime(layout) as rt:
        op = StandardizedGrgOperator(rt, k=k)
        out.update(individuals=op.n, model_variants=op.m_proj)
        log(f"operator ready: n={op.n}, M={op.m_proj}, {time.perf_counter() - t0:.1f}s")
        t = time.perf_counter()
        eig = top_eigenpairs(op, r, seed=seed)
        torch.cuda.synchronize()
        out["eig_seconds"] = time.perf_counter() - t
        out["eigenvalues"] = [float(v) for v in eig.values.cpu()]
        out["eig_rel_residual_max"] = float((eig.residuals / eig.values).max())
        basis = Basis(U=eig.vectors, lam=eig.values, trace=op.xfro2 / op.m_proj)
        log(f"top-{r} eigenpairs in {out['eig_seconds']:.1f}s: lambda_1 {out['eigenvalues'][0]:.3f}, "
            f"lambda_{r} {out['eigenvalues'][-1]:.3f}, trace {basis.trace:.1f}")

        t = time.perf_counter()
        train = simulate(op, n_train, seed=seed + 1)
        x_train, its_train, _, apps_train, _ = solve_batches(op, train, rel_tol=rel_tol, max_iter=max_iter)
        torch.cuda.synchronize()
        out["train_solve_seconds"] = time.perf_counter() - t
        out["train_mean_iterations"] = float(its_train.mean())
        log(f"{n_train} training solves in {out['train_solve_seconds']:.1f}s, mean {its_train.mean():.1f} iterations")

        t = time.perf_counter()
        # the last n_val training solves are held out: they decide whether the network is deployed
        n_val = max(k, (n_train // 8) // k * k)
        n_fit = n_train - n_val
        net = WarmStartNet(basis, seed=seed)
        history = net.fit(train.y[:, :n_fit], train.delta[:n_fit], x_train[:, :n_fit], epochs=epochs)
        torch.cuda.synchronize()
        out["train_seconds"] = time.perf_counter() - t
        out["train_loss_first_last"] = [history[0], history[-1]]
        log(f"network trained on {n_fit} solves in {out['train_seconds']:.1f}s, loss {history[0]:.4f} -> {history[-1]:.4f}")
        val = Phenotypes(y=train.y[:, n_fit:], h2=train.h2[n_fit:], delta=train.delta[n_fit:])
        _, its_net, _, _, _ = solve_batches(op, val, rel_tol=rel_tol, max_iter=max_iter,
                                            start=lambda y, d, s: net.network_start(y, d))
        _, its_cf, _, _, _ = solve_batches(op, val, rel_tol=rel_tol, max_iter=max_iter, start=lambda y, d, s:
                                           start_from_coefficients(basis, y, closed_form_coefficients(basis, y, d)))
        net.choose(float(its_net.mean()), float(its_cf.mean()))
        out["validation"] = dict(net.validation, phenotypes=int(n_val))
        log(f"validation on {n_val} held-out solves: network {its_net.mean():.2f}, closed form {its_cf.mean():.2f} "
            f"mean iterations; network deployed: {net.use_network}")

        test = simulate(op, n_test, seed=seed + 2)
        results = {}
        starts = {
            "zero": None,
            "closed_form": lambda y, d, s: start_from_coefficients(basis, y, closed_form_coefficients(basis, y, d)),
            "network": lambda y, d, s: net.start(y, d),        # as deployed (the closed form if validation rejected it)
            "network_raw": lambda y, d, s: net.network_start(y, d),
            # the least-squares best start of the same form, fitted to the zero-start solution:
            # the bound any predictor of those coefficients can reach on this data
            "best_in_span": lambda y, d, s: start_from_coefficients(
                basis, y, targets(basis, y, xs["zero"][:, s:s + y.shape[1]])),
        }
        xs = {}
        for name, start in starts.items():
            t = time.perf_counter()
            x, its, rels, apps, batch_its = solve_batches(op, test, rel_tol=rel_tol, max_iter=max_iter, start=start)
            torch.cuda.synchronize()
            xs[name] = x
            results[name] = {"mean_iterations": float(its.mean()), "max_iterations": int(its.max()),
                             "batch_iterations": batch_its, "operator_applications": apps,
                             "max_rel_residual": float(rels.max()), "seconds": time.perf_counter() - t,
                             "unconverged": int((its < 0).sum())}
            log(f"test {name}: mean {its.mean():.1f} iterations, {apps} operator applications")
        for name in ("closed_form", "network", "network_raw", "best_in_span"):
            diff = torch.linalg.norm(xs[name] - xs["zero"], dim=0) / torch.linalg.norm(xs["zero"], dim=0)
            results[name]["max_rel_diff_vs_zero_start"] = float(diff.max())
            results[name]["iterations_saved"] = 1.0 - results[name]["mean_iterations"] / results["zero"]["mean_iterations"]
            results[name]["applications_saved"] = 1.0 - results[name]["operator_applications"] / results["zero"]["operator_applications"]
        out["test"] = results
    out["seconds_total"] = time.perf_counter() - t0
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--workload", default="c5")
    ap.add_argument("--artifact-dir", default="/tmp/grg_bench")
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--r", type=int, default=20)
    ap.add_argument("--n-train", type=int, default=2048)
    ap.add_argument("--n-test", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--rel-tol", type=float, default=1e-6)
    ap.add_argument("--fp64", action="store_true")
    args = ap.parse_args()

    from pygrgl_spmv.synth import WORKLOADS, write_artifact

    cfg = WORKLOADS[args.workload]
    d = Path(args.artifact_dir)
    d.mkdir(parents=True, exist_ok=True)
    art = d / f"{args.workload}-seed{cfg.seed}.grg_spmv"
    if not art.exists():
        log(f"generating {args.workload} into {art}")
        write_artifact(cfg, art)
    res = run([art], k=args.k, r=args.r, n_train=args.n_train, n_test=args.n_test, rel_tol=args.rel_tol,
              dtype=np.float64 if args.fp64 else np.float32, epochs=args.epochs)
    res["workload"] = args.workload
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
