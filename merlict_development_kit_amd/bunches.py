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


class _DeviceArray:
    """A device array the library owns, seen through __cuda_array_interface__."""

    def __init__(self, ptr: int, n: int, typestr: str):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 3,
                                         "strides": None}


class BunchPlan:
    """Repeated batches of bunches through one captured hipGraph
    (mlgpu_bunch_plan): batches of up to max_bunches bunches and max_photons
    photons, each with its own config seed and photon id offset; the
    arrivals accumulate in `counts` when a binning is given."""

    def __init__(self, scene: GpuScene, max_bunches: int, max_photons: int,
                 config: PropagationConfig = PropagationConfig(), binning: Optional[ArrivalBinning] = None,
                 counts: Optional[torch.Tensor] = None, world_frame: int = 0, device="cuda"):
        self._scene = scene  # the plan reads the scene's device data
        self.max_photons = max_photons
        self.counts = None
        hist = None
        if binning is not None:
            self.counts = counts if counts is not None else binning.empty(torch.device(device))
            hist = binning.c_struct(self.counts)
        self._h = C.c_void_p()
        N.check(N.lib().mlgpu_bunch_plan_create(scene.handle, world_frame, max_bunches, max_photons,
                                                C.byref(config.c_struct()), C.byref(hist) if hist is not None else None,
                                                C.byref(self._h)), "mlgpu_bunch_plan_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().mlgpu_bunch_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def run(self, bunches: Sequence[Bunch], config: PropagationConfig, stream=None) -> int:
        """Replay the graph for this batch; returns its number of photons."""
        arr = (N.Bunch * max(1, len(bunches)))()
        for k, b in enumerate(bunches):
            arr[k].source = b.source.to_c()
            arr[k].seed = b.seed
            arr[k].num_photons = b.num_photons
        N.check(N.lib().mlgpu_bunch_plan_run(self._h, arr, len(bunches), C.byref(config.c_struct()),
                                             _stream_ptr(stream)), "mlgpu_bunch_plan_run")
        return sum(b.num_photons for b in bunches)

    def results(self, n: int) -> PropagationResult:
        """Copies of the last run's first n per-photon results (call after the
        run has finished)."""
        r = N.Results()
        N.check(N.lib().mlgpu_bunch_plan_results(self._h, C.byref(r)), "mlgpu_bunch_plan_results")
        types = {"final_frame": "<i4", "sensor_id": "<i4", "final_interaction": "|u1", "num_interactions": "|u1",
                 "error": "|u1", "time_of_flight": "<f8", "x": "<f4", "y": "<f4", "cx": "<f4", "cy": "<f4"}
        out = {k: torch.as_tensor(_DeviceArray(getattr(r, k), n, t), device="cuda").clone() for k, t in types.items()}
        out["accumulated_distance"] = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        return PropagationResult(**out)
