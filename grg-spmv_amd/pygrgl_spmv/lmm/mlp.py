"""Fully connected FP32 network on the GPU through ``libgrg_mlp.so`` (``include/grg_mlp.h``).

No reference counterpart: the network produces CG starting points for the
mixed-model solve. Forward, backward, loss and the Adam update are the
library's HIP kernels; torch only owns the device memory. There is no CPU
fallback: a missing library raises.
"""

from __future__ import annotations

import ctypes
from ctypes import c_float, c_int32, c_int64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parents[1] / "_lib" / "libgrg_mlp.so"
ABI_VERSION = 1


class MlpLibrary:
    def __init__(self, path: Path = LIB_PATH):
        if not Path(path).exists():
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build() -- the MLP kernels have no CPU fallback")
        import torch  # noqa: F401  (torch's HIP runtime first: it shares the soname libamdhip64.so.7)

        lib = ctypes.CDLL(str(path))
        lib.grg_mlp_linear_forward.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p]
        lib.grg_mlp_linear_backward.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                c_int32, c_int32, c_int32, c_int32, c_void_p]
        lib.grg_mlp_mse.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_void_p]
        lib.grg_mlp_adam.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_float, c_float, c_float, c_float,
                                     c_int32, c_void_p]
        for fn in ("grg_mlp_linear_forward", "grg_mlp_linear_backward", "grg_mlp_mse", "grg_mlp_adam"):
            getattr(lib, fn).restype = ctypes.c_int
        lib.grg_mlp_last_error.restype = ctypes.c_char_p
        lib.grg_mlp_abi_version.restype = c_int32
        if int(lib.grg_mlp_abi_version()) != ABI_VERSION:
            raise RuntimeError("libgrg_mlp.so ABI version mismatch; rebuild the extension")
        self.lib = lib

    def check(self, rc: int, fn: str) -> None:
        if int(rc) != 0:
            raise RuntimeError(f"HIP {fn} failed: status {int(rc)}: {self.lib.grg_mlp_last_error().decode()}")


_LIB: MlpLibrary | None = None


def mlp_library() -> MlpLibrary:
    global _LIB
    if _LIB is None:
        _LIB = MlpLibrary()
    return _LIB


def _ptr(t) -> c_void_p:
    return c_void_p(0 if t is None else int(t.data_ptr()))


class DenseNet:
    """``sizes[0] -> sizes[1] -> ... -> sizes[-1]`` with ReLU between layers (none at
    the output). Parameters, Adam moments and activations are FP32 device tensors."""

    def __init__(self, sizes, *, device, seed: int = 0):
        import torch

        self.sizes = [int(s) for s in sizes]
        if len(self.sizes) < 2 or min(self.sizes) < 1:
            raise ValueError("sizes must list at least an input and an output width")
        self.device = torch.device(device)
        self._lib = mlp_library()
        rng = np.random.default_rng(seed)
        self.W, self.b = [], []
        for fan_in, fan_out in zip(self.sizes[:-1], self.sizes[1:]):
            scale = np.sqrt(2.0 / fan_in)  # He initialisation for ReLU layers
            self.W.append(torch.from_numpy((rng.standard_normal((fan_out, fan_in)) * scale).astype(np.float32)).to(self.device))
            self.b.append(torch.zeros(fan_out, dtype=torch.float32, device=self.device))
        self._adam = None
        self._t = 0

    @property
    def params(self):
        out = []
        for w, b in zip(self.W, self.b):
            out += [w, b]
        return out

    def _stream(self):
        import torch

        return c_void_p(int(torch.cuda.current_stream(self.device).cuda_stream))

    def forward(self, X, *, keep: bool = False):
        """Network output for X (batch x sizes[0], FP32 device tensor). With ``keep``,
        also returns the per-layer activations needed by ``backward``."""
        import torch

        X = X.contiguous()
        acts = [X]
        s = self._stream()
        n_layers = len(self.W)
        for li, (w, b) in enumerate(zip(self.W, self.b)):
            out = torch.empty((X.shape[0], w.shape[0]), dtype=torch.float32, device=self.device)
            relu = 1 if li < n_layers - 1 else 0
            self._lib.check(self._lib.lib.grg_mlp_linear_forward(_ptr(acts[-1]), _ptr(w), _ptr(b), _ptr(out), int(X.shape[0]),
                                                                 int(w.shape[1]), int(w.shape[0]), relu, s),
                            "grg_mlp_linear_forward")
            acts.append(out)
        return (acts[-1], acts) if keep else acts[-1]

    def gradients(self, X, T, weights=None):
        """Loss and parameter gradients of the weighted squared error for a batch."""
        import torch

        Y, acts = self.forward(X, keep=True)
        s = self._stream()
        B = int(X.shape[0])
        dY = torch.empty_like(Y)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        self._lib.check(self._lib.lib.grg_mlp_mse(_ptr(Y), _ptr(T.contiguous()), _ptr(weights), _ptr(dY), _ptr(loss), B,
                                                  int(Y.shape[1]), s), "grg_mlp_mse")
        grads = [None] * (2 * len(self.W))
        n_layers = len(self.W)
        for li in range(n_layers - 1, -1, -1):
            w = self.W[li]
            relu = 1 if li < n_layers - 1 else 0
            dW = torch.empty_like(w)
            db = torch.empty_like(self.b[li])
            dX = torch.empty_like(acts[li]) if li > 0 else None
            self._lib.check(self._lib.lib.grg_mlp_linear_backward(_ptr(acts[li]), _ptr(w), _ptr(acts[li + 1]), _ptr(dY), _ptr(dX),
                                                                  _ptr(dW), _ptr(db), B, int(w.shape[1]), int(w.shape[0]),
                                                                  relu, s), "grg_mlp_linear_backward")
            grads[2 * li], grads[2 * li + 1] = dW, db
            dY = dX
        return loss, grads

    def adam_step(self, grads, *, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        import torch

        if self._adam is None:
            self._adam = [(torch.zeros_like(p), torch.zeros_like(p)) for p in self.params]
        self._t += 1
        s = self._stream()
        for p, g, (m, v) in zip(self.params, grads, self._adam):
            self._lib.check(self._lib.lib.grg_mlp_adam(_ptr(p), _ptr(g), _ptr(m), _ptr(v), int(p.numel()), float(lr), float(beta1),
                                                       float(beta2), float(eps), int(self._t), s), "grg_mlp_adam")

    def train(self, X, T, *, weights=None, epochs: int = 200, batch: int = 128, lr: float = 1e-3, seed: int = 0):
        """Minibatch Adam on (X, T); returns the per-epoch mean training loss."""
        import torch

        n = int(X.shape[0])
        rng = np.random.default_rng(seed)
        history = []
        for _ in range(int(epochs)):
            order = torch.from_numpy(rng.permutation(n)).to(self.device)
            losses = []
            for start in range(0, n, batch):
                idx = order[start:start + batch]
                loss, grads = self.gradients(X[idx], T[idx], weights)
                self.adam_step(grads, lr=lr)
                losses.append(loss)
            history.append(float(torch.stack(losses).mean().item()))
        return history


__all__ = ["DenseNet", "MlpLibrary", "mlp_library"]
