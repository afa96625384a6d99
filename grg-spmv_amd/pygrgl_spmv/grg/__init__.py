"""Runtime-bound GRG operator semantics.

Restates the operator layer of the reference's ``pygrgl_spmv/grg/__init__.py``:

* ``RuntimeRequirements`` (``:26-40``) and ``_CudaMatmulSpec`` (``:43-54``);
* requirement / init / miss validation with the same messages (``:63-116``,
  ``:376-425``, ``:454-516``);
* ``BoundGRG`` with ``prepare_matmul_cuda`` (``:541-562``) and ``matmul``
  (``:564-734``);
* the host level sweeps and ``_build_init_biases`` (``:119-253``) that artifact
  production uses to precompute the endpoint biases stored in ``.grg_spmv``
  files. They are the artifact producer's arithmetic only: the HIP runtime never
  calls them.

The reference imports the third-party native ``pygrgl`` module at import
time; here it is optional and only consulted to accept its
``TraversalDirection`` values when it happens to be installed.
"""

from __future__ import annotations

from collections import namedtuple
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from pygrgl_spmv.backends.types import Direction, InitMode, parse_direction, parse_init_mode
from pygrgl_spmv.grg.compile import CompiledOperatorState

Mutation = namedtuple("Mutation", ["position", "allele", "ref_allele", "time"])


@dataclass(frozen=True)
class RuntimeRequirements:
    """What a runtime must be able to run; sizes its arenas (reference ``:26-40``)."""

    max_k_up: int
    max_k_down: int
    need_down_miss_input: bool
    need_up_miss_output: bool
    need_init_vector: bool
    need_init_matrix: bool
    need_init_xtx: bool

    def __post_init__(self) -> None:
        if int(self.max_k_up) < 1:
            raise ValueError(f"max_k_up must be >= 1, got {self.max_k_up}")
        if int(self.max_k_down) < 1:
            raise ValueError(f"max_k_down must be >= 1, got {self.max_k_down}")


@dataclass(frozen=True)
class _CudaMatmulSpec:
    direction: Direction
    k: int
    emit_all_nodes: bool
    by_individual: bool
    init_mode: InitMode
    backend_init_mode: InitMode
    use_miss: bool
    input_cols: int
    output_cols: int
    apply_endpoint_bias: bool


def _decode_allele(data: np.ndarray, offsets: np.ndarray, idx: int) -> str:
    return bytes(np.asarray(data[int(offsets[idx]) : int(offsets[idx + 1])], dtype=np.uint8)).decode("utf-8")


def _validate_runtime_requirements(requirements, *, direction: Direction, k: int, init_mode: InitMode, uses_miss: bool) -> None:
    if direction == Direction.UP and int(k) > int(requirements.max_k_up):
        raise ValueError(f"UP runtime k={k} exceeds declared max_k_up={requirements.max_k_up}")
    if direction == Direction.DOWN and int(k) > int(requirements.max_k_down):
        raise ValueError(f"DOWN runtime k={k} exceeds declared max_k_down={requirements.max_k_down}")
    if direction == Direction.UP and uses_miss and not requirements.need_up_miss_output:
        raise ValueError("UP miss output was not declared in RuntimeRequirements")
    if direction == Direction.DOWN and uses_miss and not requirements.need_down_miss_input:
        raise ValueError("DOWN miss input was not declared in RuntimeRequirements")
    if init_mode == InitMode.VECTOR and not requirements.need_init_vector:
        raise ValueError("init vector use was not declared in RuntimeRequirements")
    if init_mode == InitMode.MATRIX and not requirements.need_init_matrix:
        raise ValueError("init matrix use was not declared in RuntimeRequirements")
    if init_mode == InitMode.XTX and not requirements.need_init_xtx:
        raise ValueError("init='xtx' use was not declared in RuntimeRequirements")


def _apply_internal_init(state, dtype, values: np.ndarray, *, init_mode: InitMode, init_payload) -> None:
    """Seed node values (internal order) for one sweep (reference ``:119-141``)."""
    if init_mode == InitMode.NONE:
        return
    if init_mode == InitMode.XTX:
        if state.coalescence_counts is None:
            raise ValueError("init_mode=xtx requires GRG coalescence counts")
        values += (2.0 * state.coalescence_counts.astype(dtype))[:, None]
    elif init_mode == InitMode.VECTOR:
        values += init_payload[None, :]
    elif init_mode == InitMode.MATRIX:
        values += init_payload
    else:
        raise ValueError(f"unknown init mode {init_mode!r}")


def _host_sweep(state, dtype, values: np.ndarray, direction: Direction) -> None:
    """In-place level sweep over ``state.A_blocks`` on the host (reference ``:158-166,195-204``)."""
    off = np.asarray(state.level_offsets, dtype=np.int64)
    height = off.size - 1
    if direction == Direction.UP:
        for d in range(1, height):
            dst = values[off[d] : off[d + 1]]
            for s, block in enumerate(state.A_blocks[d]):
                if block.nnz:
                    dst += block @ values[off[s] : off[s + 1]]
    else:
        for d in range(height - 2, -1, -1):
            dst = values[off[d] : off[d + 1]]
            for s in range(height - 1, d, -1):
                block = state.A_blocks[s][d]
                if block.nnz:
                    dst += block.T @ values[off[s] : off[s + 1]]


def _host_run_up(state, dtype, x: np.ndarray, *, init_mode: InitMode, init_payload):
    values = np.zeros((state.num_nodes, x.shape[1]), dtype=dtype)
    _apply_internal_init(state, dtype, values, init_mode=init_mode, init_payload=init_payload)
    values[: state.num_samples] += x
    _host_sweep(state, dtype, values, Direction.UP)
    if state.sel_mut.nnz:
        return np.asarray(state.sel_mut @ values, dtype=dtype)
    return np.zeros((state.num_mutations, x.shape[1]), dtype=dtype)


def _host_run_down(state, dtype, y: np.ndarray, *, init_mode: InitMode, init_payload):
    values = np.zeros((state.num_nodes, y.shape[1]), dtype=dtype)
    _apply_internal_init(state, dtype, values, init_mode=init_mode, init_payload=init_payload)
    if state.sel_mut.nnz:
        values += state.sel_mut.T @ y
    _host_sweep(state, dtype, values, Direction.DOWN)
    return values[: state.num_samples]


def _build_init_biases(compiled: CompiledOperatorState, dtype) -> None:
    """Endpoint biases an ``init`` sweep adds to the plain result (reference ``:208-253``).

    ``init_vector_{up,down}_bias`` are the outputs of a sweep with zero input
    and init vector ``[1]``; ``init_xtx_{up,down}_bias`` likewise for
    ``init='xtx'``. By linearity ``sweep(x, init=c) = sweep(x) + bias * c``.
    """
    if compiled.A_blocks is None:
        raise RuntimeError("compiled operator blocks are required to build init biases")
    dtype = np.dtype(dtype)
    zeros_up = np.zeros((compiled.num_samples, 1), dtype=dtype)
    zeros_down = np.zeros((compiled.num_mutations, 1), dtype=dtype)
    one = np.ones(1, dtype=dtype)
    compiled.init_vector_up_bias = _host_run_up(compiled, dtype, zeros_up, init_mode=InitMode.VECTOR, init_payload=one)[:, 0].copy()
    compiled.init_vector_down_bias = _host_run_down(compiled, dtype, zeros_down, init_mode=InitMode.VECTOR, init_payload=one)[:, 0].copy()
    compiled.init_xtx_up_bias = None
    compiled.init_xtx_down_bias = None
    if compiled.coalescence_counts is not None:
        compiled.init_xtx_up_bias = _host_run_up(compiled, dtype, zeros_up, init_mode=InitMode.XTX, init_payload=None)[:, 0].copy()
        compiled.init_xtx_down_bias = _host_run_down(compiled, dtype, zeros_down, init_mode=InitMode.XTX, init_payload=None)[:, 0].copy()


def _traversal_direction_types():
    try:  # optional third-party; only used to accept its direction enum
        import pygrgl  # type: ignore

        return pygrgl.TraversalDirection
    except Exception:
        return None


class BoundGRG:
    """One GRG bound to an entered runtime (reference ``:256-734``)."""

    def __init__(self, runtime, artifact_index: int, state: CompiledOperatorState, artifact_path: Path, device_id: int | None = None):
        self._runtime = runtime
        self._artifact_index = int(artifact_index)
        self._state = state
        self._artifact_path = Path(artifact_path)
        self._dtype = np.dtype(runtime.layout.dtype)
        self._device_id = device_id

    device = property(lambda self: self._device_id)
    shape = property(lambda self: (self.num_samples, self.num_mutations))
    artifact_path = property(lambda self: self._artifact_path)
    num_samples = property(lambda self: int(self._state.num_samples))
    num_individuals = property(lambda self: int(self._state.num_individuals))
    num_mutations = property(lambda self: int(self._state.num_mutations))
    ploidy = property(lambda self: int(self._state.ploidy))
    num_nodes = property(lambda self: int(self._state.num_nodes))
    num_edges = property(lambda self: int(self._state.num_edges))
    has_missing_data = property(lambda self: bool(self._state.has_missing_data))
    level_offsets = property(lambda self: self._state.level_offsets)
    node_perm = property(lambda self: self._state.node_perm)
    inv_node_perm = property(lambda self: self._state.inv_node_perm)
    sel_mut = property(lambda self: self._state.sel_mut)
    sel_miss = property(lambda self: self._state.sel_miss)
    sample_to_individual = property(lambda self: self._state.sample_to_individual)
    coalescence_counts = property(lambda self: self._state.coalescence_counts)
    init_vector_up_bias = property(lambda self: self._state.init_vector_up_bias)
    init_vector_down_bias = property(lambda self: self._state.init_vector_down_bias)
    init_xtx_up_bias = property(lambda self: self._state.init_xtx_up_bias)
    init_xtx_down_bias = property(lambda self: self._state.init_xtx_down_bias)

    def get_mutation_by_id(self, mutation_id: int):
        idx = int(mutation_id)
        if idx < 0 or idx >= self.num_mutations:
            raise IndexError(f"mutation id out of range: {mutation_id}")
        st = self._state
        return Mutation(
            float(st.mutation_positions[idx]),
            _decode_allele(st.mutation_alleles, st.mutation_allele_offsets, idx),
            _decode_allele(st.mutation_ref_alleles, st.mutation_ref_allele_offsets, idx),
            float(st.mutation_times[idx]),
        )

    def _parse_direction(self, direction) -> Direction:
        if isinstance(direction, (str, Direction)):
            return parse_direction(direction)
        td = _traversal_direction_types()
        if td is not None:
            if direction == td.UP:
                return Direction.UP
            if direction == td.DOWN:
                return Direction.DOWN
        raise ValueError(
            f"unknown direction: {direction!r}. Expected 'up', 'down', "
            "pygrgl.TraversalDirection.UP, or pygrgl.TraversalDirection.DOWN"
        )

    def _parse_init(self, init, rows: int, input_dtype, *, reorder_matrix: bool):
        if init is None:
            return InitMode.NONE, None
        if isinstance(init, str):
            if init != "xtx":
                raise ValueError(f"unexpected init value: {init}")
            if self.coalescence_counts is None:
                raise ValueError(
                    "init='xtx' requires per-node coalescence counts in the GRG. "
                    "This artifact was loaded without coalescence counts."
                )
            return InitMode.XTX, None
        if not isinstance(init, np.ndarray):
            raise TypeError(f"init must be None, 'xtx', or a numpy.ndarray, got {type(init).__name__}")
        if init.dtype != input_dtype:
            raise TypeError(f"the init matrix must match the dtype of the input matrix. Got: {init.dtype}")
        if init.ndim == 1:
            if init.shape[0] != rows:
                raise ValueError("if init has a single dimension, it must match the number of rows in the input matrix")
            return InitMode.VECTOR, init.astype(self._dtype, order="C", copy=False)
        if init.ndim == 2:
            if init.shape != (rows, self.num_nodes):
                raise ValueError(f"if init is a matrix, it must match the dimensions ({rows}, {self.num_nodes})")
            if reorder_matrix:
                return InitMode.MATRIX, init[:, self.node_perm].T.astype(self._dtype, order="C", copy=False)
            return InitMode.MATRIX, init.astype(self._dtype, order="C", copy=False)
        raise ValueError("init must be None, 'xtx', a vector, or a matrix")

    def _validate_miss(self, miss, rows: int, direction: Direction, input_dtype) -> np.ndarray:
        if not isinstance(miss, np.ndarray):
            raise TypeError(f'the "miss" input must be a numpy.ndarray. Got: {type(miss).__name__}')
        if miss.dtype != input_dtype:
            raise TypeError(f'the "miss" input must match the dtype of the input matrix. Got: {miss.dtype}')
        if miss.ndim != 2:
            raise ValueError(f'"miss" must be a two-dimension numpy array (matrix). ndim={miss.ndim}')
        if miss.shape[0] != rows:
            raise ValueError(f'"miss" has {miss.shape[0]} rows, but must match the input/output matrices ({rows})')
        if miss.shape[1] != self.num_mutations:
            if direction == Direction.DOWN:
                raise ValueError(f'the "miss" matrix must match the number of columns in the input matrix. Got: {miss.shape[1]}')
            raise ValueError(f'the "miss" matrix must match the number of columns in the output matrix. Got: {miss.shape[1]}')
        return miss

    def _apply_endpoint_init_bias(self, result_internal, *, direction: Direction, init_mode: InitMode, init_payload) -> None:
        if init_mode == InitMode.XTX:
            bias = self.init_xtx_up_bias if direction == Direction.UP else self.init_xtx_down_bias
            result_internal += bias[:, None]
        elif init_mode == InitMode.VECTOR:
            bias = self.init_vector_up_bias if direction == Direction.UP else self.init_vector_down_bias
            result_internal += bias[:, None] * init_payload[None, :]

    def _finish_endpoint_output(self, result_internal):
        return result_internal.T.astype(self._dtype, copy=False)

    def _finish_node_output(self, node_values_internal):
        values = np.asarray(node_values_internal, dtype=self._dtype, order="C")
        return values[self.inv_node_perm].T.astype(self._dtype, copy=False)

    def _prepare_cuda_spec(self, *, direction, k: int, emit_all_nodes: bool, by_individual: bool, init_mode, use_miss: bool) -> _CudaMatmulSpec:
        d = self._parse_direction(direction)
        mode = parse_init_mode(init_mode)
        rows = int(k)
        if rows < 1:
            raise ValueError(f"k must be >= 1, got {k}")
        if emit_all_nodes and use_miss:
            raise RuntimeError('the "use_miss" parameter cannot be mixed with "emit_all_nodes=True"')
        if mode != InitMode.NONE and use_miss:
            raise ValueError('the "use_miss" parameter cannot be mixed with a non-none init_mode')
        if mode == InitMode.XTX and self.coalescence_counts is None:
            raise ValueError(
                "init_mode='xtx' requires per-node coalescence counts in the GRG. "
                "This artifact was loaded without coalescence counts."
            )
        _validate_runtime_requirements(self._runtime.layout.requirements, direction=d, k=rows, init_mode=mode, uses_miss=bool(use_miss))
        if d == Direction.UP:
            input_cols = self.num_individuals if by_individual else self.num_samples
        else:
            input_cols = self.num_mutations
        if emit_all_nodes:
            output_cols = self.num_nodes
        elif d == Direction.UP:
            output_cols = self.num_mutations
        else:
            output_cols = self.num_individuals if by_individual else self.num_samples
        return _CudaMatmulSpec(
            direction=d,
            k=rows,
            emit_all_nodes=bool(emit_all_nodes),
            by_individual=bool(by_individual),
            init_mode=mode,
            backend_init_mode=mode if (emit_all_nodes or mode == InitMode.MATRIX) else InitMode.NONE,
            use_miss=bool(use_miss),
            input_cols=int(input_cols),
            output_cols=int(output_cols),
            apply_endpoint_bias=bool((not emit_all_nodes) and mode in (InitMode.VECTOR, InitMode.XTX)),
        )

    def _copy_host_to_cuda(self, dst, src: np.ndarray) -> None:
        import torch

        arr = np.asarray(src, dtype=self._dtype, order="C")
        stream_fn = getattr(self._runtime, "_torch_caller_stream", None)
        if stream_fn is None:
            dst.copy_(torch.from_numpy(arr), non_blocking=False)
            return
        with torch.cuda.device(dst.device), torch.cuda.stream(stream_fn()):
            dst.copy_(torch.from_numpy(arr), non_blocking=False)

    def _copy_cuda_to_host(self, src) -> np.ndarray:
        import torch

        stream_fn = getattr(self._runtime, "_torch_caller_stream", None)
        if stream_fn is None:
            return src.detach().cpu().numpy().astype(self._dtype, copy=False).copy()
        with torch.cuda.device(src.device), torch.cuda.stream(stream_fn()):
            host = src.detach().cpu()
        return host.numpy().astype(self._dtype, copy=False).copy()

    def prepare_matmul_cuda(self, *, direction, k: int, emit_all_nodes: bool = False, by_individual: bool = False, init_mode="none", use_miss: bool = False):
        """A reusable device op with ``input``/``output``/... tensor views (reference ``:541-562``)."""
        prepare = getattr(self._runtime, "_prepare_matmul_cuda", None)
        if prepare is None:
            raise NotImplementedError("prepare_matmul_cuda() is implemented only for CUDA backends")
        spec = self._prepare_cuda_spec(
            direction=direction,
            k=int(k),
            emit_all_nodes=bool(emit_all_nodes),
            by_individual=bool(by_individual),
            init_mode=init_mode,
            use_miss=bool(use_miss),
        )
        return prepare(self, spec)

    def matmul(self, input, direction, emit_all_nodes: bool = False, by_individual: bool = False, init=None, miss=None):
        """Eager ``input @ X`` (up) or ``input @ X.T`` (down) on host arrays (reference ``:564-734``)."""
        if not isinstance(input, np.ndarray):
            raise TypeError(f"matmul() requires input to be a numpy.ndarray, got {type(input).__name__}")
        if input.ndim != 2:
            raise ValueError("matmul() only supports two-dimensional numpy arrays as input.")
        rows, cols = input.shape
        if rows == 0 or cols == 0:
            raise ValueError("matmul() requires non-zero dimensions.")
        d = self._parse_direction(direction)
        if d == Direction.UP:
            expected = self.num_individuals if by_individual else self.num_samples
        else:
            expected = self.num_mutations
        if cols != expected:
            if d == Direction.UP:
                raise ValueError("input matrix has wrong number of columns for UP direction (numSamples or numIndividuals depending on by_individual)")
            raise ValueError("input matrix has wrong number of columns for DOWN direction (numMutations)")
        if emit_all_nodes and miss is not None:
            raise RuntimeError('the "miss" parameter cannot be mixed with the "emit_all_nodes" parameter')
        if init is not None and miss is not None:
            raise ValueError('the "miss" parameter cannot be mixed with the "init" parameter')

        if getattr(self._runtime, "_prepare_matmul_cuda", None) is not None:
            init_mode, payload = self._parse_init(init, rows, input.dtype, reorder_matrix=False)
            spec = self._prepare_cuda_spec(
                direction=d, k=rows, emit_all_nodes=bool(emit_all_nodes), by_individual=bool(by_individual),
                init_mode=init_mode, use_miss=miss is not None,
            )
            miss_matrix = None if miss is None else self._validate_miss(miss, rows, d, input.dtype)
            with self._runtime._call_scope():
                with self.prepare_matmul_cuda(
                    direction=spec.direction, k=spec.k, emit_all_nodes=spec.emit_all_nodes,
                    by_individual=spec.by_individual, init_mode=spec.init_mode, use_miss=spec.use_miss,
                ) as op:
                    self._copy_host_to_cuda(op.input, input)
                    if d == Direction.DOWN and miss_matrix is not None:
                        self._copy_host_to_cuda(op.miss_input, miss_matrix)
                    if init_mode == InitMode.VECTOR:
                        self._copy_host_to_cuda(op.init_vector, payload)
                    elif init_mode == InitMode.MATRIX:
                        self._copy_host_to_cuda(op.init_matrix, payload)
                    run = getattr(op, "_call_prelocked", None)
                    if run is None:
                        raise RuntimeError("prepared CUDA op is missing _call_prelocked()")
                    run()
                    result = self._copy_cuda_to_host(op.output)
                    if d == Direction.UP and miss_matrix is not None:
                        miss_matrix += self._copy_cuda_to_host(op.miss_output).astype(miss_matrix.dtype, copy=False)
                    return result

        # Host runtimes expose ``_run`` (reference ``:636-734``).
        init_mode, payload = self._parse_init(init, rows, input.dtype, reorder_matrix=True)
        _validate_runtime_requirements(self._runtime.layout.requirements, direction=d, k=rows, init_mode=init_mode, uses_miss=miss is not None)
        if emit_all_nodes:
            b_mode, b_payload = init_mode, payload
        elif init_mode in (InitMode.VECTOR, InitMode.XTX):
            b_mode, b_payload = InitMode.NONE, None
        else:
            b_mode, b_payload = init_mode, payload
        x_internal = input.astype(self._dtype, order="C", copy=False).T
        with self._runtime._call_scope():
            if d == Direction.UP:
                if by_individual:
                    x_internal = x_internal[self.sample_to_individual]
                if emit_all_nodes:
                    nodes = self._runtime._run(self._artifact_index, d, x_internal, miss=None, init_mode=b_mode,
                                               init_payload=b_payload, need_miss_output=False, emit_all_nodes=True)
                    return self._finish_node_output(nodes)
                miss_out = None if miss is None else self._validate_miss(miss, rows, d, input.dtype)
                res, miss_res = self._runtime._run(self._artifact_index, d, x_internal, miss=None, init_mode=b_mode,
                                                   init_payload=b_payload, need_miss_output=miss_out is not None, emit_all_nodes=False)
                if init_mode != InitMode.NONE:
                    self._apply_endpoint_init_bias(res, direction=d, init_mode=init_mode, init_payload=payload)
                if miss_out is not None and miss_res is not None:
                    miss_out += miss_res.T.astype(miss_out.dtype, copy=False)
                return self._finish_endpoint_output(res)
            if emit_all_nodes:
                nodes = self._runtime._run(self._artifact_index, d, x_internal, miss=None, init_mode=b_mode,
                                           init_payload=b_payload, need_miss_output=False, emit_all_nodes=True)
                return self._finish_node_output(nodes)
            miss_internal = None
            if miss is not None:
                miss_internal = self._validate_miss(miss, rows, d, input.dtype).T.astype(self._dtype, order="C", copy=False)
            res = self._runtime._run(self._artifact_index, d, x_internal, miss=miss_internal, init_mode=b_mode,
                                     init_payload=b_payload, need_miss_output=False, emit_all_nodes=False)
            if init_mode != InitMode.NONE:
                self._apply_endpoint_init_bias(res, direction=d, init_mode=init_mode, init_payload=payload)
            if by_individual:
                acc = np.zeros((self.num_individuals, rows), dtype=self._dtype)
                np.add.at(acc, self.sample_to_individual, res)
                res = acc
            return self._finish_endpoint_output(res)


from pygrgl_spmv.grg.convert import convert  # noqa: E402  (convert imports this module's helpers lazily)

__all__ = ["BoundGRG", "Mutation", "RuntimeRequirements", "convert"]
