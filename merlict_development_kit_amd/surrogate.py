"""The learned surrogate of the propagator (mlgpu_surrogate): an MLP from a
photon's normalised (support x, y, direction cx, cy, wavelength) to its
survival logit and an 8-component Gaussian mixture over sensor-plane
position and arrival time, trained on the GPU by the library's own bf16
matrix-core kernels on photons the plain propagator produced."""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N

NUM_OUTPUTS = 57
NUM_COMPONENTS = 8


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream(stream) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Surrogate:
    def __init__(self, seed: int = 0):
        self._h = C.c_void_p()
        N.check(N.lib().mlgpu_surrogate_create(seed, C.byref(self._h)), "mlgpu_surrogate_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().mlgpu_surrogate_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @staticmethod
    def num_params() -> int:
        return int(N.lib().mlgpu_surrogate_num_params())

    def params(self, device="cuda") -> torch.Tensor:
        out = torch.empty(self.num_params(), dtype=torch.float32, device=device)
        N.check(N.lib().mlgpu_surrogate_get_params(self._h, _ptr(out), _stream(None)), "mlgpu_surrogate_get_params")
        torch.cuda.synchronize()
        return out

    def set_params(self, params: torch.Tensor) -> None:
        p = params.to(dtype=torch.float32).contiguous()
        if p.numel() != self.num_params():
            raise ValueError("wrong number of parameters")
        N.check(N.lib().mlgpu_surrogate_set_params(self._h, _ptr(p), _stream(None)), "mlgpu_surrogate_set_params")
        torch.cuda.synchronize()

    @staticmethod
    def _check(features: torch.Tensor, targets: torch.Tensor = None):
        if features.dtype != torch.float32 or features.dim() != 2 or features.shape[1] != 5 \
                or not features.is_contiguous() or features.device.type != "cuda":
            raise ValueError("features must be a contiguous (n, 5) float32 tensor on the GPU")
        if targets is not None and (targets.dtype != torch.float32 or tuple(targets.shape) != (len(features), 4)
                                    or not targets.is_contiguous() or targets.device != features.device):
            raise ValueError("targets must be a contiguous (n, 4) float32 tensor on the features' device")

    def forward(self, features: torch.Tensor, stream=None) -> torch.Tensor:
        self._check(features)
        out = torch.empty((len(features), NUM_OUTPUTS), dtype=torch.float32, device=features.device)
        N.check(N.lib().mlgpu_surrogate_forward(self._h, _ptr(features), len(features), _ptr(out), _stream(stream)),
                "mlgpu_surrogate_forward")
        return out

    def gradients(self, features: torch.Tensor, targets: torch.Tensor, stream=None):
        """(mean loss, gradient of the flat parameters)."""
        self._check(features, targets)
        g = torch.empty(self.num_params(), dtype=torch.float32, device=features.device)
        loss = torch.empty(1, dtype=torch.float32, device=features.device)
        N.check(N.lib().mlgpu_surrogate_gradients(self._h, _ptr(features), _ptr(targets), len(features), _ptr(g),
                                                  _ptr(loss), _stream(stream)), "mlgpu_surrogate_gradients")
        return loss / len(features), g

    def train_step(self, features: torch.Tensor, targets: torch.Tensor, learning_rate: float,
                   loss: torch.Tensor = None, stream=None) -> None:
        """One Adam step; `loss` (1-element float32 on the GPU) receives the
        summed loss of the batch before the step."""
        self._check(features, targets)
        N.check(N.lib().mlgpu_surrogate_train_step(self._h, _ptr(features), _ptr(targets), len(features),
                                                   float(learning_rate), _ptr(loss) if loss is not None else None,
                                                   _stream(stream)), "mlgpu_surrogate_train_step")
