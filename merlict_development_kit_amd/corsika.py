"""CORSIKA photon bunches to photons on the GPU (mlgpu_corsika_photons).

Replaces merlict::EventIoPhotonFactory (merlict_corsika/PhotonFactory.cpp:8-148)
as the reference's apps apply it to every bunch of an event
(merlict_portal_plenoscope/apps/plenoscope_propagation.cpp:196-203).  A bunch
is 8 float32: x, y [cm], cx, cy, arrival time on the ground [ns], emission
height [cm], weight, wavelength [nm].
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N
from .propagation import PhotonBatch


class BadPhotonWeight(ValueError):
    """A bunch weight outside [0, 16] (EventIoPhotonFactory::BadPhotonWeight)."""


def photons_from_bunches(bunches: torch.Tensor, seed: int = 0, id_offset: int = 0, stream=None):
    """Photons of the bunches (an (n, 8) float32 tensor on the GPU), in bunch
    order, and the bunch index of each (int32).  Bunch b's weight draw is
    Philox stream 5 of (seed, id_offset + b).  Raises BadPhotonWeight like
    the reference."""
    if bunches.dtype != torch.float32 or bunches.dim() != 2 or bunches.shape[1] != 8 \
            or not bunches.is_contiguous() or bunches.device.type != "cuda":
        raise ValueError("bunches must be a contiguous (n, 8) float32 tensor on the GPU")
    n = len(bunches)
    dev = bunches.device
    # every photon count is at most floor(w) + 1
    w = bunches[:, 6].double()
    ok = (w >= 0) & (w <= 16)
    capacity = int((torch.floor(torch.where(ok, w, torch.zeros_like(w))).sum() + n).item()) if n else 0
    arrays = [torch.empty(capacity, dtype=torch.float64, device=dev) for _ in range(7)]
    truth = torch.empty(capacity, dtype=torch.int32, device=dev)
    buf = N.PhotonBuffers(*[a.data_ptr() if capacity else None for a in arrays])
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    m = C.c_int64()
    rc = N.lib().mlgpu_corsika_photons(C.c_void_p(bunches.data_ptr()), n, seed, id_offset, C.byref(buf),
                                       C.c_void_p(truth.data_ptr() if capacity else None), capacity, C.byref(m),
                                       C.c_void_p(st.cuda_stream))
    if rc != N.MLGPU_OK:
        msg = N.lib().mlgpu_last_error().decode(errors="replace")
        if "weight" in msg:
            raise BadPhotonWeight(msg)
        N.check(rc, "mlgpu_corsika_photons")
    k = m.value
    return PhotonBatch(*[a[:k] for a in arrays]), truth[:k]
