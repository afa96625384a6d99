"""Loading and graph-replay layer over the HIP backend.

Restates the convenience layer of the reference (``pygrgl_spmv/adaptor.py``)
for ``pygrgl_spmv.backends.hip``:

* ``load_grg_spmv_single`` / ``load_grg_spmv_multi`` plan and enter one
  ``HipRuntime`` per device for a set of ``.grg_spmv`` files (reference
  ``adaptor.py:983-1090``);
* with ``capture=True`` every GRG comes back as a ``CapturedBoundGRG``: the
  prepared ops named by a ``RunConfigs`` are captured once into hipGraphs
  (``torch.cuda.CUDAGraph`` on ROCm) and ``matmul`` replays them (reference
  ``adaptor.py:29-420,850-965``), with the reference's input contract: k up to
  the captured k (zero-padded before replay, truncated after), int32 inputs
  accepted by FP64 captures, ``init`` / ``miss`` payloads staged with the input;
* ``CaptureSpec`` / ``RunConfigs`` and the ``make_runconfig_*`` factories name the
  graphs a workload needs (reference ``adaptor.py:428-680``).

Native mode takes and returns torch tensors on the capture device (the
reference's native mode uses CuPy arrays, which this build does not use); copy
mode takes and returns numpy arrays. The reference's CPU MKL backend config has
no counterpart here: this package ships the HIP backend only.
"""

from __future__ import annotations

import contextlib
import threading
from collections import defaultdict
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from pygrgl_spmv.grg import RuntimeRequirements

_VALID_INIT_MODES = ("none", "vector", "matrix", "xtx")


# ---------------------------------------------------------------------------
# capture specifications (reference adaptor.py:428-451)
# ---------------------------------------------------------------------------

@dataclass(frozen=True)
class CaptureSpec:
    """One ``prepare_matmul_cuda`` configuration to capture as a graph."""

    direction: str
    k: int
    by_individual: bool = False
    init_mode: str = "none"   # "none" | "vector" | "matrix" | "xtx"
    use_miss: bool = False


@dataclass(frozen=True)
class RunConfigs:
    """Requirements for the layout plus the graphs to capture (``capture_ops``; when
    empty, one up and one down graph at the requirements' widths)."""

    req: RuntimeRequirements
    capture_ops: tuple = field(default_factory=tuple)


@dataclass(frozen=True)
class HipBackendConfig:
    """HIP backend options (the counterpart of the reference's cuSPARSE config,
    ``adaptor.py:466-487``).

    ``device``: a device id for every file, or ``{file stem: {"device": id}}``.
    ``allow_residency``: resident blocks when they fit; with False every block
    streams through a two-slot ring under ``vram_budget_mb`` (which must be > 0).
    ``capture``: return ``CapturedBoundGRG`` objects with pre-captured graphs.
    ``native``: captured ``matmul`` takes and returns device tensors.
    """

    device: object = 0
    allow_residency: bool = True
    vram_budget_mb: int = 0
    capture: bool = False
    native: bool = False


def make_backend_hip(device=0, allow_residency=True, vram_budget_mb=0, capture=False, native=False, **kwargs) -> HipBackendConfig:
    """Validated ``HipBackendConfig`` (argument checks as reference ``adaptor.py:505-537``)."""
    if kwargs:
        raise TypeError(f"make_backend_hip() got unexpected keyword arguments: {sorted(kwargs)}")
    if not isinstance(device, (int, dict)):
        raise TypeError(f"device must be int or dict, got {type(device).__name__}")
    if isinstance(device, int) and device < 0:
        raise ValueError(f"device must be >= 0, got {device}")
    if not isinstance(vram_budget_mb, int):
        raise TypeError(f"vram_budget_mb must be int, got {type(vram_budget_mb).__name__}")
    if vram_budget_mb < 0:
        raise ValueError(f"vram_budget_mb must be >= 0, got {vram_budget_mb}")
    if not allow_residency and vram_budget_mb == 0:
        raise ValueError("vram_budget_mb must be > 0 in streaming mode (allow_residency=False)")
    if native and not capture:
        raise ValueError("native=True needs capture=True (native I/O is a property of captured matmul)")
    return HipBackendConfig(device=device, allow_residency=bool(allow_residency), vram_budget_mb=int(vram_budget_mb),
                            capture=bool(capture), native=bool(native))


def _requirements(k: int, *, miss_in=False, miss_out=False, vector=False, matrix=False, xtx=False) -> RuntimeRequirements:
    return RuntimeRequirements(max_k_up=int(k), max_k_down=int(k), need_down_miss_input=miss_in, need_up_miss_output=miss_out,
                               need_init_vector=vector, need_init_matrix=matrix, need_init_xtx=xtx)


def make_runconfig_kernel(direction, k, force_spmm=False) -> RunConfigs:
    """One plain graph of the benchmarked direction (reference ``adaptor.py:540-560``);
    ``force_spmm`` captures k = 1 at width 2 (k = 1 callers are padded)."""
    if direction not in ("up", "down"):
        raise ValueError(f"direction must be 'up' or 'down', got {direction!r}")
    width = 2 if (force_spmm and int(k) == 1) else int(k)
    return RunConfigs(req=_requirements(width), capture_ops=(CaptureSpec(direction, width),))


def make_runconfig_pca(force_spmm=False, maxk=1, **kwargs) -> RunConfigs:
    """Graphs of a PCA workload (reference ``adaptor.py:563-596``): allele counts up
    (with and without the missingness output) and the by-individual up / down pair."""
    if kwargs:
        raise TypeError(f"make_runconfig_pca() got unexpected keyword arguments: {sorted(kwargs)}")
    width = (2 if force_spmm else 1) if maxk == 1 else int(maxk)
    return RunConfigs(
        req=_requirements(width, miss_out=True, vector=True),
        capture_ops=(CaptureSpec("up", width), CaptureSpec("up", width, use_miss=True),
                     CaptureSpec("up", width, by_individual=True), CaptureSpec("down", width, by_individual=True)),
    )


def make_runconfig_bolt(force_spmm=False, **kwargs) -> RunConfigs:
    """Graphs of a BOLT-LMM-style workload (reference ``adaptor.py:599-635``): plain,
    XTX-initialised and missing-data sweeps, by individual and by sample."""
    if kwargs:
        raise TypeError(f"make_runconfig_bolt() got unexpected keyword arguments: {sorted(kwargs)}")
    width = 2 if force_spmm else 1
    ops = [CaptureSpec("up", width, by_individual=True), CaptureSpec("up", width, by_individual=True, init_mode="xtx"),
           CaptureSpec("down", width, by_individual=True), CaptureSpec("up", width, by_individual=True, use_miss=True),
           CaptureSpec("down", width, by_individual=True, use_miss=True)]
    for d in ("up", "down"):
        ops.append(CaptureSpec(d, width))
    for d in ("up", "down"):
        ops.append(CaptureSpec(d, width, use_miss=True))
    return RunConfigs(req=_requirements(width, miss_in=True, miss_out=True, xtx=True), capture_ops=tuple(ops))


def make_runconfig_gwas(force_spmm=False, maxk=1, sample_variance=True, **kwargs) -> RunConfigs:
    """Graphs of an association-test workload (reference ``adaptor.py:638-680``)."""
    if kwargs:
        raise TypeError(f"make_runconfig_gwas() got unexpected keyword arguments: {sorted(kwargs)}")
    if not isinstance(maxk, int) or maxk < 1:
        raise ValueError(f"maxk must be an int >= 1, got {maxk!r}")
    width = (2 if force_spmm else 1) if maxk == 1 else maxk
    ops = [CaptureSpec("up", width, use_miss=True), CaptureSpec("up", width, by_individual=True),
           CaptureSpec("up", width, by_individual=True, use_miss=True)]
    if sample_variance:
        ops.append(CaptureSpec("up", width, init_mode="xtx"))
    return RunConfigs(req=_requirements(width, miss_out=True, xtx=bool(sample_variance)), capture_ops=tuple(ops))


def _bare_req(req) -> RuntimeRequirements:
    return req.req if isinstance(req, RunConfigs) else req


def _validate_capture_spec(spec: CaptureSpec, req: RuntimeRequirements) -> None:
    """As reference ``adaptor.py:828-847``."""
    if spec.init_mode not in _VALID_INIT_MODES:
        raise ValueError(f"CaptureSpec.init_mode {spec.init_mode!r} invalid; expected one of {sorted(_VALID_INIT_MODES)}")
    if spec.init_mode != "none" and spec.use_miss:
        raise ValueError(f"CaptureSpec direction={spec.direction!r}: init_mode and use_miss are mutually exclusive")
    if spec.use_miss:
        if spec.direction == "up" and not req.need_up_miss_output:
            raise ValueError("CaptureSpec use_miss=True for 'up' but need_up_miss_output is False")
        if spec.direction == "down" and not req.need_down_miss_input:
            raise ValueError("CaptureSpec use_miss=True for 'down' but need_down_miss_input is False")
    if spec.init_mode == "vector" and not req.need_init_vector:
        raise ValueError("CaptureSpec init_mode='vector' but need_init_vector is False")
    if spec.init_mode == "matrix" and not req.need_init_matrix:
        raise ValueError("CaptureSpec init_mode='matrix' but need_init_matrix is False")
    if spec.init_mode == "xtx" and not req.need_init_xtx:
        raise ValueError("CaptureSpec init_mode='xtx' but need_init_xtx is False")


# ---------------------------------------------------------------------------
# graph-replay GRG (reference adaptor.py:29-420)
# ---------------------------------------------------------------------------

class CapturedBoundGRG:
    """A ``BoundGRG`` whose ``matmul`` replays pre-captured graphs.

    Every other attribute is the wrapped GRG's. ``matmul`` stages the caller's
    input (and ``init`` / ``miss``) into the captured buffers on the capture
    stream, replays, and returns a fresh result: numpy in copy mode, a device
    tensor in native mode. It returns only once the result (and the in-place
    ``miss`` accumulation of an up sweep) is complete. GRGs that share a device
    share one lock, so their replays never interleave.
    """

    def __init__(self, grg, prepared_ops, graphs, src_tensors, init_tensors, miss_tensors, capture_stream, *, native=False,
                 device_lock=None):
        self._grg = grg
        self._prepared_ops = prepared_ops
        self._graphs = graphs
        self._src_tensors = src_tensors
        self._init_tensors = init_tensors
        self._miss_tensors = miss_tensors
        self._capture_stream = capture_stream
        self._native = bool(native)
        self._device_lock = device_lock if device_lock is not None else contextlib.nullcontext()

    @property
    def native(self) -> bool:
        return self._native

    @property
    def captured_keys(self) -> tuple:
        return tuple(self._graphs)

    def __getattr__(self, name):
        if name == "_grg":
            raise AttributeError(name)
        return getattr(self._grg, name)

    def _key(self, direction, by_individual, init_mode, use_miss):
        key = (direction, bool(by_individual), init_mode, bool(use_miss))
        if key in self._graphs:
            return key
        # as the reference: fall back to the plain graph of that direction and grouping
        plain = (direction, bool(by_individual), "none", False)
        if plain in self._graphs and init_mode == "none" and not use_miss:
            return plain
        raise KeyError(f"no captured graph for direction={direction!r}, by_individual={by_individual}, "
                       f"init_mode={init_mode!r}, use_miss={use_miss}; captured: {sorted(self._graphs)}")

    @staticmethod
    def _np_dtype(d) -> np.dtype:
        if isinstance(d, np.dtype):
            return d
        if getattr(type(d), "__module__", "").startswith("torch"):
            import torch

            table = {torch.float32: np.float32, torch.float64: np.float64, torch.int32: np.int32}
            if d not in table:
                raise TypeError(f"unsupported torch dtype {d!r}")
            return np.dtype(table[d])
        return np.dtype(d)

    def _validate(self, arr, *, name, role, expected_shape, expected_dtype, expected_device, exact_k=False):
        """Type, device, dtype and shape checks of one payload (reference ``adaptor.py:84-146``);
        ``exact_k``: axis 0 must equal the expected k (init / miss payloads use the input's k)."""
        if arr is None:
            raise ValueError(f"matmul(): {name} ({role}) is required but got None")
        want = self._np_dtype(expected_dtype)
        if expected_device is not None:
            import torch

            if not isinstance(arr, torch.Tensor):
                raise TypeError(f"matmul(): {name} ({role}) expected a torch.Tensor on {expected_device} (native mode), "
                                f"got {type(arr).__name__}")
            if arr.device != expected_device:
                raise TypeError(f"matmul(): {name} ({role}) expected device {expected_device}, got {arr.device}")
            got = self._np_dtype(arr.dtype)
        else:
            if not isinstance(arr, np.ndarray):
                raise TypeError(f"matmul(): {name} ({role}) expected numpy.ndarray (copy mode), got {type(arr).__name__}")
            got = arr.dtype
        if got != want and not (want == np.dtype(np.float64) and got == np.dtype(np.int32)):
            raise TypeError(f"matmul(): {name} ({role}) dtype mismatch: expected {want}, got {got}")
        shape, exp = tuple(arr.shape), tuple(expected_shape)
        if len(shape) != len(exp) or shape[0] > exp[0] or (exact_k and shape[0] != exp[0]) or shape[1:] != exp[1:]:
            bound = "==" if exact_k else "<="
            raise ValueError(f"matmul(): {name} ({role}) shape mismatch: expected k {bound} {exp[0]} with trailing dims "
                             f"{exp[1:]}, got {shape}")

    def matmul(self, input, direction, emit_all_nodes=False, by_individual=False, init=None, miss=None):
        """Replay the captured graph of this configuration (reference ``adaptor.py:163-425``).

        ``input`` is ``(k, input_cols)`` with k at most the captured k; ``init`` is
        None, ``"xtx"``, a ``(k,)`` vector or a ``(k, num_nodes)`` matrix; ``miss`` is
        the down sweep's ``(k, num_mutations)`` input, or the up sweep's output
        accumulator of that shape, updated in place (``+=``)."""
        import torch

        if emit_all_nodes:
            raise ValueError("emit_all_nodes=True is not supported by CapturedBoundGRG.matmul()")
        with self._device_lock:
            if init is None:
                init_mode, payload = "none", None
            elif isinstance(init, str):
                if init != "xtx":
                    raise ValueError(f"unexpected init value: {init!r}")
                init_mode, payload = "xtx", None
            elif getattr(init, "ndim", 2) == 1:
                init_mode, payload = "vector", init
            else:
                init_mode, payload = "matrix", init
            use_miss = miss is not None
            key = self._key(direction, by_individual, init_mode, use_miss)
            src = self._src_tensors[key]
            device = src.device if self._native else None
            int_input = src.dtype == torch.float64 and self._np_dtype(getattr(input, "dtype", np.float64)) == np.dtype(np.int32)
            self._validate(input, name="input", role="input", expected_shape=tuple(src.shape), expected_dtype=src.dtype,
                           expected_device=device)
            k = int(input.shape[0])
            full = int(src.shape[0])

            def like_input(shape):
                return (k,) + tuple(shape[1:])

            if key in self._init_tensors:
                self._validate(payload, name="init", role="init", expected_shape=like_input(self._init_tensors[key].shape),
                               expected_dtype=src.dtype, expected_device=device, exact_k=True)
            op = self._prepared_ops[key]
            if key in self._miss_tensors:
                self._validate(miss, name="miss", role="miss-in", expected_shape=like_input(self._miss_tensors[key].shape),
                               expected_dtype=src.dtype, expected_device=device, exact_k=True)
            elif use_miss:
                self._validate(miss, name="miss", role="miss-out", expected_shape=like_input(op.miss_output.shape),
                               expected_dtype=src.dtype, expected_device=device, exact_k=True)

            def as_tensor(arr):
                t = arr if isinstance(arr, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(arr))
                return t.to(src.dtype) if t.dtype != src.dtype else t

            def stage(dst, arr):
                # zero-pad a narrower k up to the captured k
                t = as_tensor(arr)
                if k < full:
                    dst[:k].copy_(t)
                    dst[k:].zero_()
                else:
                    dst.copy_(t)

            cap = self._capture_stream
            if self._native:
                # the caller's producer work on its current stream comes first
                cap.wait_stream(torch.cuda.current_stream(src.device))
            with torch.cuda.stream(cap):
                stage(src, input)
                if key in self._init_tensors:
                    stage(self._init_tensors[key], payload)
                if key in self._miss_tensors:
                    stage(self._miss_tensors[key], miss)
                self._graphs[key].replay()
                if self._native:
                    result = op.output[:k, :].clone()
                    if int_input:
                        result = result.to(torch.int32)
                    if use_miss and key not in self._miss_tensors:
                        miss += op.miss_output[:k, :].to(miss.dtype)
                else:
                    result = op.output[:k, :].cpu().numpy().copy()
                    if int_input:
                        result = result.astype(np.int32)
                    if use_miss and key not in self._miss_tensors:
                        miss += op.miss_output[:k, :].cpu().numpy().astype(miss.dtype, copy=False)
            cap.synchronize()
            return result


def _capture_grg(grg, capture_stream, req, cfg, stack, device_lock=None) -> CapturedBoundGRG:
    """Prepare every op of ``req.capture_ops`` (entered into ``stack``), size one
    shared staging buffer for inputs and one each for init / miss payloads (ops of
    one GRG never run at once), and capture one graph per op: stage copies from
    the shared buffers into the op's views, then the op (reference
    ``adaptor.py:850-965``)."""
    import torch

    if isinstance(req, RunConfigs) and req.capture_ops:
        specs = tuple(req.capture_ops)
    else:
        bare = _bare_req(req)
        specs = (CaptureSpec("up", bare.max_k_up), CaptureSpec("down", bare.max_k_down))
    bare = _bare_req(req)
    ops = {}
    for spec in specs:
        _validate_capture_spec(spec, bare)
        key = (spec.direction, bool(spec.by_individual), spec.init_mode, bool(spec.use_miss))
        if key in ops:
            raise ValueError(f"Duplicate CaptureSpec key {key!r} in capture_ops")
        op = stack.enter_context(grg.prepare_matmul_cuda(direction=spec.direction, k=spec.k, by_individual=spec.by_individual,
                                                         init_mode=spec.init_mode, use_miss=spec.use_miss))
        if not op.capturable:
            raise ValueError(f"CaptureSpec {key!r}: this sweep streams blocks from host memory and is not capturable; "
                             "use a resident layout (allow_residency=True with a sufficient budget)")
        ops[key] = (spec, op)

    def numel(spec, op):
        if spec.init_mode == "vector":
            return op.init_vector.numel()
        if spec.init_mode == "matrix":
            return op.init_matrix.numel()
        return 0

    def miss_numel(spec, op):
        if spec.use_miss:
            return op.miss_input.numel() if spec.direction == "down" else op.miss_output.numel()
        return 0

    first = next(iter(ops.values()))[1]
    dtype, device = first.input.dtype, first.input.device
    shared_src = torch.zeros(max(op.input.numel() for _, op in ops.values()), dtype=dtype, device=device)
    n_init = max(numel(s, o) for s, o in ops.values())
    n_miss = max(miss_numel(s, o) for s, o in ops.values())
    shared_init = torch.zeros(n_init, dtype=dtype, device=device) if n_init else None
    shared_miss = torch.zeros(n_miss, dtype=dtype, device=device) if n_miss else None

    graphs, srcs, inits, misses = {}, {}, {}, {}
    for key, (spec, op) in ops.items():
        k, cols = op.input.shape
        # column-major staging views, laid out as the op's own (k, cols) views
        src = torch.as_strided(shared_src, size=(k, cols), stride=(1, k))
        staged = [(op.input, src)]
        if spec.init_mode == "vector":
            inits[key] = shared_init[:k]
            staged.append((op.init_vector, inits[key]))
        elif spec.init_mode == "matrix":
            inits[key] = torch.as_strided(shared_init, size=(k, op.init_matrix.shape[1]), stride=(1, k))
            staged.append((op.init_matrix, inits[key]))
        elif spec.use_miss and spec.direction == "down":
            misses[key] = torch.as_strided(shared_miss, size=(k, op.miss_input.shape[1]), stride=(1, k))
            staged.append((op.miss_input, misses[key]))
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=capture_stream):
            for dst, s in staged:
                dst.copy_(s)
            op()
        graphs[key] = graph
        srcs[key] = src
    return CapturedBoundGRG(grg, {k: o for k, (_, o) in ops.items()}, graphs, srcs, inits, misses, capture_stream,
                            native=cfg.native, device_lock=device_lock)


# ---------------------------------------------------------------------------
# loading (reference adaptor.py:983-1090)
# ---------------------------------------------------------------------------

def _validate_filename(path: Path) -> None:
    if path.suffix != ".grg_spmv":
        raise ValueError(f"expected a .grg_spmv file, got {path}")
    if not path.exists():
        raise FileNotFoundError(path)


def _validate_dtype(dtype) -> np.dtype:
    dtype = np.dtype(dtype)
    if dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError(f"dtype must be float32 or float64, got {dtype}")
    return dtype


def _resolve_device(cfg: HipBackendConfig, path: Path) -> int:
    from pygrgl_spmv.backends.hip.ffi import hip_library

    if isinstance(cfg.device, int):
        device = cfg.device
    else:
        if path.stem not in cfg.device:
            raise KeyError(f"file stem {path.stem!r} not found in device config; available stems: {sorted(cfg.device)}")
        entry = cfg.device[path.stem]
        if "device" not in entry:
            raise KeyError(f"device config for {path.stem!r} is missing 'device' key; got: {entry!r}")
        device = int(entry["device"])
    n = hip_library().device_count()
    if device >= n:
        raise RuntimeError(f"HIP device {device} is not available; system has {n} device(s)")
    return device


def _load_hip(paths, cfg: HipBackendConfig, req, stack, dtype):
    import torch

    from pygrgl_spmv.backends.hip import HipPlanPair, HipRuntime, plan_hip_layout

    groups = defaultdict(list)
    for i, path in enumerate(paths):
        groups[_resolve_device(cfg, path)].append((i, path))
    out = {}
    for device, members in groups.items():
        lock = threading.Lock()
        with torch.cuda.device(device):
            stream = torch.cuda.Stream(device=device) if cfg.capture else 0
            if cfg.allow_residency:
                budget, ring = 0, 0
            else:
                budget, ring = cfg.vram_budget_mb * 1024 * 1024, 2
            layout = plan_hip_layout(artifacts=[p for _, p in members], pair=HipPlanPair.default(), dtype=dtype,
                                     requirements=_bare_req(req), vram_budget_bytes=budget, ring_buffer_size=ring,
                                     allow_residency=cfg.allow_residency, device=device, stream=stream)
            runtime = stack.enter_context(HipRuntime(layout))
            for (index, _), grg in zip(members, runtime.grgs):
                if cfg.capture:
                    grg = _capture_grg(grg, stream, req, cfg, stack, device_lock=lock)
                out[index] = grg
    return [out[i] for i in range(len(paths))]


def load_grg_spmv_multi(filenames, backend, req, stack, dtype=np.float64):
    """Load ``.grg_spmv`` files onto their devices; one runtime per device, entered
    into ``stack`` (which must outlive the returned GRGs). ``req`` is a
    ``RunConfigs`` or a ``RuntimeRequirements``."""
    paths = [Path(f) for f in filenames]
    if not paths:
        raise ValueError("filenames must be non-empty")
    for p in paths:
        _validate_filename(p)
    if not isinstance(backend, HipBackendConfig):
        raise TypeError(f"backend must be HipBackendConfig (see make_backend_hip), got {type(backend).__name__}")
    if not isinstance(req, (RunConfigs, RuntimeRequirements)):
        raise TypeError(f"req must be RunConfigs or RuntimeRequirements, got {type(req).__name__}")
    if not isinstance(stack, contextlib.ExitStack):
        raise TypeError(f"stack must be contextlib.ExitStack, got {type(stack).__name__}")
    return _load_hip(paths, backend, req, stack, _validate_dtype(dtype))


def load_grg_spmv_single(filename, backend, req, stack, dtype=np.float64):
    """``load_grg_spmv_multi`` of one file."""
    return load_grg_spmv_multi([filename], backend, req, stack, dtype)[0]


__all__ = [
    "CaptureSpec", "CapturedBoundGRG", "HipBackendConfig", "RunConfigs", "load_grg_spmv_multi", "load_grg_spmv_single",
    "make_backend_hip", "make_runconfig_bolt", "make_runconfig_gwas", "make_runconfig_kernel", "make_runconfig_pca",
]
