"""Loading and graph-replay layer (``pygrgl_spmv.adaptor``).

Counterparts in the reference: ``pygrgl_spmv/adaptor.py`` (captured ``matmul``
contract: k padding and truncation, int32 coercion, init / miss staging, copy and
native I/O, multi-file loading) and its use by the PCA / BOLT scripts. The GPU
cases compare every replay with the CPU oracle on the same inputs (FP64 within
1e-12, exact for integer inputs); the CPU cases check the configuration
contracts without a GPU.
"""

from __future__ import annotations

import contextlib

import numpy as np
import pytest

from oracle import grg_oracle as O
from tests._builders import write_synthetic_artifact

FP64_TOL = 1e-12


def _cfg(seed, samples=240, mutations=1500, ploidy=2, missing=0.05):
    from pygrgl_spmv.synth import SynthConfig

    return SynthConfig(num_samples=samples, num_mutations=mutations, seed=seed, ploidy=ploidy, num_trees=100,
                       keep_fraction=0.3, window_fraction=0.1, num_segments=2, missing_fraction=missing, missing_pool=5)


@pytest.fixture(scope="module")
def art_a(artifact_dir):
    return write_synthetic_artifact(artifact_dir, "adaptor-a", _cfg(611))


@pytest.fixture(scope="module")
def art_b(artifact_dir):
    return write_synthetic_artifact(artifact_dir, "adaptor-b", _cfg(612, samples=160, mutations=900, ploidy=1, missing=0.0))


# ---------------------------------------------------------------------------
# configuration contracts (CPU)
# ---------------------------------------------------------------------------

def test_make_backend_hip_validation():
    from pygrgl_spmv.adaptor import HipBackendConfig, make_backend_hip

    assert make_backend_hip() == HipBackendConfig(device=0, allow_residency=True, vram_budget_mb=0, capture=False, native=False)
    with pytest.raises(TypeError, match="unexpected keyword"):
        make_backend_hip(n_threads=4)
    with pytest.raises(TypeError, match="device must be int or dict"):
        make_backend_hip(device="0")
    with pytest.raises(ValueError, match="device must be >= 0"):
        make_backend_hip(device=-1)
    with pytest.raises(ValueError, match="vram_budget_mb must be >= 0"):
        make_backend_hip(vram_budget_mb=-5)
    with pytest.raises(ValueError, match="streaming mode"):
        make_backend_hip(allow_residency=False)
    with pytest.raises(ValueError, match="needs capture=True"):
        make_backend_hip(native=True)
    assert make_backend_hip(allow_residency=False, vram_budget_mb=64).vram_budget_mb == 64


def test_runconfig_factories():
    from pygrgl_spmv.adaptor import CaptureSpec, make_runconfig_bolt, make_runconfig_gwas, make_runconfig_kernel, make_runconfig_pca

    rc = make_runconfig_kernel("down", 1, force_spmm=True)
    assert rc.capture_ops == (CaptureSpec("down", 2),) and rc.req.max_k_down == 2
    with pytest.raises(ValueError, match="direction"):
        make_runconfig_kernel("left", 4)
    pca = make_runconfig_pca(maxk=8)
    assert pca.req.need_init_vector and pca.req.need_up_miss_output and pca.req.max_k_up == 8
    assert CaptureSpec("down", 8, by_individual=True) in pca.capture_ops and len(pca.capture_ops) == 4
    bolt = make_runconfig_bolt()
    assert bolt.req.need_init_xtx and bolt.req.need_down_miss_input and len(bolt.capture_ops) == 9
    assert len({(s.direction, s.by_individual, s.init_mode, s.use_miss) for s in bolt.capture_ops}) == 9
    gwas = make_runconfig_gwas(maxk=3, sample_variance=False)
    assert not gwas.req.need_init_xtx and all(s.init_mode == "none" for s in gwas.capture_ops)
    with pytest.raises(ValueError, match="maxk"):
        make_runconfig_gwas(maxk=0)
    with pytest.raises(TypeError, match="unexpected keyword"):
        make_runconfig_pca(width=3)


def test_capture_spec_validation():
    from pygrgl_spmv.adaptor import CaptureSpec, _validate_capture_spec
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=2, max_k_down=2, need_down_miss_input=False, need_up_miss_output=True,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=True)
    _validate_capture_spec(CaptureSpec("up", 2, use_miss=True), req)
    _validate_capture_spec(CaptureSpec("up", 2, init_mode="xtx"), req)
    for spec, msg in ((CaptureSpec("up", 2, init_mode="diag"), "invalid"),
                      (CaptureSpec("up", 2, init_mode="xtx", use_miss=True), "mutually exclusive"),
                      (CaptureSpec("down", 2, use_miss=True), "need_down_miss_input"),
                      (CaptureSpec("up", 2, init_mode="vector"), "need_init_vector"),
                      (CaptureSpec("up", 2, init_mode="matrix"), "need_init_matrix")):
        with pytest.raises(ValueError, match=msg):
            _validate_capture_spec(spec, req)


def test_load_argument_validation(tmp_path, art_a):
    from pygrgl_spmv.adaptor import load_grg_spmv_multi, make_backend_hip
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=1, max_k_down=1, need_down_miss_input=False, need_up_miss_output=False,
                              need_init_vector=False, need_init_matrix=False, need_init_xtx=False)
    be = make_backend_hip()
    with contextlib.ExitStack() as stack:
        with pytest.raises(ValueError, match="non-empty"):
            load_grg_spmv_multi([], be, req, stack)
        with pytest.raises(ValueError, match=".grg_spmv"):
            load_grg_spmv_multi([tmp_path / "x.npz"], be, req, stack)
        with pytest.raises(FileNotFoundError):
            load_grg_spmv_multi([tmp_path / "missing.grg_spmv"], be, req, stack)
        with pytest.raises(TypeError, match="HipBackendConfig"):
            load_grg_spmv_multi([art_a], object(), req, stack)
        with pytest.raises(TypeError, match="RunConfigs or RuntimeRequirements"):
            load_grg_spmv_multi([art_a], be, {"max_k_up": 1}, stack)
        with pytest.raises(ValueError, match="float32 or float64"):
            load_grg_spmv_multi([art_a], be, req, stack, dtype=np.int32)
    with pytest.raises(TypeError, match="ExitStack"):
        load_grg_spmv_multi([art_a], be, req, None)


# ---------------------------------------------------------------------------
# captured replays on the GPU
# ---------------------------------------------------------------------------

def _bolt_like(k):
    from pygrgl_spmv.adaptor import CaptureSpec, RunConfigs
    from pygrgl_spmv.grg import RuntimeRequirements

    req = RuntimeRequirements(max_k_up=k, max_k_down=k, need_down_miss_input=True, need_up_miss_output=True,
                              need_init_vector=True, need_init_matrix=True, need_init_xtx=True)
    ops = (CaptureSpec("up", k), CaptureSpec("down", k), CaptureSpec("down", k, use_miss=True),
           CaptureSpec("up", k, use_miss=True), CaptureSpec("up", k, by_individual=True, init_mode="xtx"),
           CaptureSpec("up", k, init_mode="vector"), CaptureSpec("down", k, by_individual=True, init_mode="matrix"))
    return RunConfigs(req=req, capture_ops=ops)


@pytest.mark.gpu
def test_captured_replays_match_oracle(art_a):
    """Every captured configuration, replayed several times at the captured k and at a
    smaller k (zero-padded, truncated), with other eager work on the arena between
    replays, matches the oracle; int32 input to an FP64 capture comes back int32 and
    exact."""
    from pygrgl_spmv.adaptor import CapturedBoundGRG, load_grg_spmv_single, make_backend_hip

    g = O.load_artifact(art_a)
    k = 4
    rng = np.random.default_rng(6100)
    with contextlib.ExitStack() as stack:
        grg = load_grg_spmv_single(art_a, make_backend_hip(capture=True), _bolt_like(k), stack)
        assert isinstance(grg, CapturedBoundGRG) and grg.num_samples == g.num_samples
        for rep in range(3):
            for kk in (k, 2):
                x = rng.standard_normal((kk, g.num_samples))
                y = rng.standard_normal((kk, g.num_mutations))
                got = grg.matmul(x, "up")
                assert O.normwise_rel_err(got, O.matmul(g, x, "up")) <= FP64_TOL and got.shape == (kk, g.num_mutations)
                assert O.normwise_rel_err(grg.matmul(y, "down"), O.matmul(g, y, "down")) <= FP64_TOL
                miss = rng.standard_normal((kk, g.num_mutations))
                assert O.normwise_rel_err(grg.matmul(y, "down", miss=miss), O.matmul(g, y, "down", miss=miss)) <= FP64_TOL
                acc, acc_ref = np.ones((kk, g.num_mutations)), np.ones((kk, g.num_mutations))
                got = grg.matmul(x, "up", miss=acc)
                assert O.normwise_rel_err(got, O.matmul(g, x, "up", miss=acc_ref)) <= FP64_TOL
                assert O.normwise_rel_err(acc, acc_ref) <= FP64_TOL
                xi = rng.standard_normal((kk, g.num_individuals))
                assert O.normwise_rel_err(grg.matmul(xi, "up", by_individual=True, init="xtx"),
                                          O.matmul(g, xi, "up", by_individual=True, init="xtx")) <= FP64_TOL
                c = rng.standard_normal(kk)
                assert O.normwise_rel_err(grg.matmul(x, "up", init=c), O.matmul(g, x, "up", init=c)) <= FP64_TOL
                m0 = rng.standard_normal((kk, g.num_nodes))
                assert O.normwise_rel_err(grg.matmul(y, "down", by_individual=True, init=m0),
                                          O.matmul(g, y, "down", by_individual=True, init=m0)) <= FP64_TOL
            # eager work on the same arena between replays
            assert callable(grg.prepare_matmul_cuda)  # other attributes are the wrapped GRG's
            ones = np.ones((1, g.num_samples))
            np.testing.assert_array_equal(grg._grg.matmul(ones, "up"), O.matmul(g, ones, "up"))
        xint = rng.integers(-3, 4, size=(3, g.num_samples)).astype(np.int32)
        got = grg.matmul(xint, "up")
        assert got.dtype == np.int32
        np.testing.assert_array_equal(got, O.matmul(g, xint.astype(np.float64), "up").astype(np.int32))


@pytest.mark.gpu
def test_captured_native_mode_and_errors(art_a):
    """Native mode takes and returns device tensors and agrees with copy mode; the
    input contract rejects wrong types, dtypes, widths and missing payloads."""
    import torch

    from pygrgl_spmv.adaptor import load_grg_spmv_single, make_backend_hip

    g = O.load_artifact(art_a)
    rng = np.random.default_rng(6200)
    x = rng.standard_normal((3, g.num_samples))
    y = rng.standard_normal((3, g.num_mutations))
    with contextlib.ExitStack() as stack:
        native = load_grg_spmv_single(art_a, make_backend_hip(capture=True, native=True), _bolt_like(4), stack)
        dev = torch.device("cuda", 0)
        xt = torch.from_numpy(x).to(dev)
        got = native.matmul(xt, "up")
        assert isinstance(got, torch.Tensor) and got.device == dev
        assert O.normwise_rel_err(got.cpu().numpy(), O.matmul(g, x, "up")) <= FP64_TOL
        acc = torch.zeros((3, g.num_mutations), dtype=torch.float64, device=dev)
        acc_ref = np.zeros((3, g.num_mutations))
        native.matmul(xt, "up", miss=acc)
        O.matmul(g, x, "up", miss=acc_ref)
        assert O.normwise_rel_err(acc.cpu().numpy(), acc_ref) <= FP64_TOL
        with pytest.raises(TypeError, match="torch.Tensor"):
            native.matmul(x, "up")
        copy = load_grg_spmv_single(art_a, make_backend_hip(capture=True), _bolt_like(4), stack)
        with pytest.raises(TypeError, match="numpy.ndarray"):
            copy.matmul(xt, "up")
        with pytest.raises(TypeError, match="dtype mismatch"):
            copy.matmul(x.astype(np.float32), "up")
        with pytest.raises(ValueError, match="shape mismatch"):
            copy.matmul(rng.standard_normal((5, g.num_samples)), "up")
        with pytest.raises(ValueError, match="shape mismatch"):
            copy.matmul(rng.standard_normal((2, g.num_samples + 1)), "up")
        with pytest.raises(ValueError, match="init .* shape mismatch"):
            copy.matmul(y, "down", by_individual=True, init=rng.standard_normal((2, g.num_nodes)))
        with pytest.raises(ValueError, match="miss .* shape mismatch"):
            copy.matmul(y, "down", miss=rng.standard_normal((4, g.num_mutations)))
        with pytest.raises(KeyError, match="no captured graph"):
            copy.matmul(x, "up", by_individual=True)
        with pytest.raises(ValueError, match="emit_all_nodes"):
            copy.matmul(x, "up", emit_all_nodes=True)


@pytest.mark.gpu
def test_multi_file_load(art_a, art_b):
    """Two files on one device share one runtime; captured and plain GRGs both match
    the oracle, in file order."""
    from pygrgl_spmv.adaptor import load_grg_spmv_multi, make_backend_hip, make_runconfig_kernel
    from pygrgl_spmv.grg import BoundGRG

    gs = [O.load_artifact(p) for p in (art_a, art_b)]
    rng = np.random.default_rng(6300)
    with contextlib.ExitStack() as stack:
        cap = load_grg_spmv_multi([art_a, art_b], make_backend_hip(capture=True), make_runconfig_kernel("down", 2), stack)
        plain = load_grg_spmv_multi([art_b, art_a], make_backend_hip(), make_runconfig_kernel("up", 3).req, stack)
        assert isinstance(plain[0], BoundGRG)
        for grg, g in zip(cap, gs):
            y = rng.standard_normal((2, g.num_mutations))
            assert O.normwise_rel_err(grg.matmul(y, "down"), O.matmul(g, y, "down")) <= FP64_TOL
        for grg, g in zip(plain, gs[::-1]):
            x = rng.standard_normal((3, g.num_samples))
            assert O.normwise_rel_err(grg.matmul(x, "up"), O.matmul(g, x, "up")) <= FP64_TOL
