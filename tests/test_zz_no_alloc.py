"""``grg_hip_gather_batch_static``: the sweep launches without a caller workspace (GPU).

The entry point runs the same launches as ``grg_hip_gather_batch`` but keeps the
split-row partials in the library's static per-GPU array. Two calls in a row,
the second on a smaller input (another GRG, fewer columns), each give bit for
bit the output of the workspace entry point on the same descriptors (and the
oracle's within 1e-12); a launch whose partials exceed the array is refused
before anything runs.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from oracle import grg_oracle as O
from tests._builders import build_hip_layout, full_requirements, write_synthetic_artifact

pytestmark = pytest.mark.gpu


def _cfg(seed, samples, mutations):
    from pygrgl_spmv.synth import SynthConfig

    return SynthConfig(num_samples=samples, num_mutations=mutations, seed=seed, ploidy=2, num_trees=120,
                       keep_fraction=0.3, window_fraction=0.1, num_segments=3, missing_fraction=0.05, missing_pool=6)


def _run_batches(rt, op, entry) -> None:
    import torch

    stream = torch.cuda.current_stream(rt.device)
    for kind, arr, n in op._steps:
        assert kind == "batch"
        entry(arr, n, int(stream.cuda_stream))
    torch.cuda.synchronize(rt.device)


def _split_launches(op) -> int:
    from pygrgl_spmv.backends.hip import ffi

    return sum(1 for _, arr, n in op._steps for i in range(n) if arr[i].flags & ffi.FLAG_PARTIALS)


def test_static_scratch_entry_point_matches_workspace_entry_point(artifact_dir):
    import torch

    from pygrgl_spmv.backends.hip import HipRuntime, ffi

    big = write_synthetic_artifact(artifact_dir, "noalloc-big", _cfg(611, 600, 5000))
    small = write_synthetic_artifact(artifact_dir, "noalloc-small", _cfg(612, 200, 900))
    oracles = [O.load_artifact(big), O.load_artifact(small)]
    layout = build_hip_layout([big, small], requirements=full_requirements(max_k_up=8, max_k_down=8))
    rng = np.random.default_rng(6100)
    # small split limits: many rows of every level are summed from split-row partials
    with HipRuntime(layout, split_min=12, split_chunk=5, autotune=False) as rt:
        lib = rt._lib
        calls = [(0, "down", 8, True), (1, "up", 3, False)]   # the second call: a smaller GRG, fewer columns
        for gi, direction, k, use_miss in calls:
            grg, g = rt.grgs[gi], oracles[gi]
            op = grg.prepare_matmul_cuda(direction=direction, k=k, use_miss=use_miss and direction == "down")
            assert _split_launches(op) > 0
            cols = g.num_samples if direction == "up" else g.num_mutations
            x = rng.standard_normal((k, cols))
            miss = rng.standard_normal((k, g.num_mutations)) if op._spec.use_miss else None

            def set_inputs():
                # the output view shares the arena with the input view, so inputs are set before every run
                op.input.copy_(torch.from_numpy(x))
                if miss is not None:
                    op.miss_input.copy_(torch.from_numpy(miss))

            set_inputs()
            _run_batches(rt, op, lib.gather_batch)
            with_workspace = op.output.detach().cpu().clone()
            rt._partials.fill_(float("nan"))   # the workspace entry point's partials are not read
            set_inputs()
            _run_batches(rt, op, lib.gather_batch_static)
            static = op.output.detach().cpu().clone()
            assert torch.equal(static, with_workspace), (gi, direction, k)
            want = O.matmul(g, x, direction, miss=miss) if direction == "down" else O.matmul(g, x, direction)
            assert O.normwise_rel_err(static.numpy(), want) <= 1e-12


def test_static_scratch_entry_point_refuses_oversized_partials(artifact_dir):
    from pygrgl_spmv.backends.hip import HipRuntime, ffi

    art = write_synthetic_artifact(artifact_dir, "noalloc-big", _cfg(611, 600, 5000))
    layout = build_hip_layout([art], requirements=full_requirements(max_k_up=8, max_k_down=8))
    with HipRuntime(layout, split_min=12, split_chunk=5, autotune=False) as rt:
        (grg,) = rt.grgs
        op = grg.prepare_matmul_cuda(direction="down", k=8)
        part = next(arr[i] for _, arr, n in op._steps for i in range(n) if arr[i].flags & ffi.FLAG_PARTIALS)
        one = (ffi.GrgHipGatherArgs * 1)()
        ctypes.pointer(one[0])[0] = part
        one[0].split_items = ffi.STATIC_SCRATCH_DOUBLES // one[0].k + 1
        with pytest.raises(RuntimeError, match="more than the static scratch holds"):
            rt._lib.gather_batch_static(one, 1, 0)
