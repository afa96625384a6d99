"""Device / stream argument parsing for the HIP backend.

Same contract as the reference's ``backends/_cuda_stream.py:108-174``: a device
is a non-bool int ordinal that must be in range; a stream is a non-negative int
handle (0 = default stream) or an object implementing ``__cuda_stream__``
(returning a handle or ``(0, handle)``). torch HIP streams are accepted too
(their ``cuda_stream`` attribute is the ``hipStream_t``). Device count and a
stream's device come from ``libgrg_hip.so`` (hipGetDeviceCount /
hipStreamGetDevice), not from a CUDA runtime.
"""

from __future__ import annotations

from numbers import Integral


class HipStreamToken:
    """Wrap a raw ``hipStream_t`` in the ``__cuda_stream__`` protocol."""

    def __init__(self, ptr: int) -> None:
        self._ptr = int(ptr)

    def __cuda_stream__(self) -> tuple[int, int]:
        return 0, self._ptr


def _lib():
    from pygrgl_spmv.backends.hip.ffi import hip_library

    return hip_library()


def parse_hip_device(device: object, *, device_count=None) -> int:
    if isinstance(device, bool) or not isinstance(device, Integral):
        raise TypeError(f"HIP device must be a non-bool int ordinal, got {type(device).__name__}")
    device_id = int(device)
    if device_id < 0:
        raise ValueError(f"HIP device ordinal must be non-negative, got {device_id}")
    count = _lib().device_count() if device_count is None else int(device_count())
    if device_id >= count:
        raise ValueError(f"HIP device ordinal {device_id} is out of range for {count} visible HIP devices")
    return device_id


def parse_hip_stream(stream: object) -> tuple[int, object | None]:
    """``(hipStream_t, owner_ref)`` from an int, a ``__cuda_stream__`` object or a torch stream."""
    if isinstance(stream, bool):
        raise TypeError("HIP stream must be a non-bool int handle or an object implementing __cuda_stream__()")
    if isinstance(stream, Integral):
        ptr = int(stream)
        if ptr < 0:
            raise ValueError(f"HIP stream handle must be non-negative, got {ptr}")
        return ptr, None
    method = getattr(stream, "__cuda_stream__", None)
    if method is None:
        raw = getattr(stream, "cuda_stream", None)  # torch.cuda.Stream on ROCm
        if isinstance(raw, Integral) and not isinstance(raw, bool):
            if int(raw) < 0:
                raise ValueError(f"HIP stream handle must be non-negative, got {int(raw)}")
            return int(raw), stream
        raise TypeError("HIP stream must be a non-bool int handle or an object implementing __cuda_stream__()")
    token = method()
    if isinstance(token, bool):
        raise TypeError("__cuda_stream__() must return an int handle or a (version, handle) pair")
    if isinstance(token, Integral):
        ptr = int(token)
    elif isinstance(token, (tuple, list)):
        if len(token) != 2:
            raise TypeError("__cuda_stream__() tuple/list return must be exactly (version, handle)")
        version, ptr = token
        if int(version) != 0:
            raise ValueError(f"Unsupported CUDA stream protocol version {int(version)}; expected 0")
        ptr = int(ptr)
    else:
        raise TypeError("__cuda_stream__() must return an int handle or a (version, handle) pair")
    if ptr < 0:
        raise ValueError(f"HIP stream handle must be non-negative, got {ptr}")
    return ptr, stream


def hip_stream_device(ptr: int) -> int:
    if int(ptr) == 0:
        raise ValueError("hipStreamGetDevice requires a non-null HIP stream handle")
    return _lib().stream_device(int(ptr))


__all__ = ["HipStreamToken", "hip_stream_device", "parse_hip_device", "parse_hip_stream"]
