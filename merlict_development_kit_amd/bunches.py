"""Many photon bunches per call (mlgpu_propagate_bunches): drawn, propagated
and binned on the device, each bunch with its own source pose and seed."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import torch

from . import _native as N
from .arrivals import ArrivalBinning
from .propagation import GpuScene, PropagationConfig, PropagationResult, _stream_ptr
from .sources import Source


@dataclass(frozen=True)
class Bunch:
    source: Source
    seed: int
    num_photons: int


def propagate_bunches(scene: GpuScene, bunches: Sequence[Bunch], config: PropagationConfig = PropagationConfig(),
                      binning: Optional[ArrivalBinning] = None, counts: Optional[torch.Tensor] = None,
                      want_results: bool = False, world_frame: int = 0, device="cuda",
                      stream=None) -> Tuple[Optional[torch.Tensor], Optional[PropagationResult]]:
    """Returns (counts, results): counts accumulate the arrivals when a
    binning is given (allocated zeroed when None), results are the per-photon
    results of the concatenated bunches when asked for."""
    device = torch.device(device)
    arr = (N.Bunch * len(bunches))()
    for k, b in enumerate(bunches):
        arr[k].source = b.source.to_c()
        arr[k].seed = b.seed
        arr[k].num_photons = b.num_photons
    total = sum(b.num_photons for b in bunches)
    res = PropagationResult.empty(total, device) if want_results else None
    hist = None
    if binning is not None:
        if counts is None:
            counts = binning.empty(device)
        hist = binning.c_struct(counts)
    cfg = config.c_struct()
    rs = res.c_struct() if res is not None else None
    N.check(N.lib().mlgpu_propagate_bunches(scene.handle, world_frame, arr, len(bunches), C.byref(cfg),
                                            C.byref(hist) if hist is not None else None,
                                            C.byref(rs) if rs is not None else None,
                                            _stream_ptr(stream)), "mlgpu_propagate_bunches")
    return counts, res


def generate_bunches(bunches: Sequence[Bunch], device="cuda", stream=None):
    """The photons of the bunches, concatenated in bunch order
    (mlgpu_generate_bunches): photon j of a bunch is photon j of its source
    under its seed."""
    from .propagation import PhotonBatch
    device = torch.device(device)
    arr = (N.Bunch * len(bunches))()
    for k, b in enumerate(bunches):
        arr[k].source = b.source.to_c()
        arr[k].seed = b.seed
        arr[k].num_photons = b.num_photons
    total = sum(b.num_photons for b in bunches)
    arrays = [torch.empty(total, dtype=torch.float64, device=device) for _ in range(7)]
    buf = N.PhotonBuffers(*[a.data_ptr() if total else None for a in arrays])
    N.check(N.lib().mlgpu_generate_bunches(arr, len(bunches), C.byref(buf), total, _stream_ptr(stream)),
            "mlgpu_generate_bunches")
    return PhotonBatch(*arrays)
