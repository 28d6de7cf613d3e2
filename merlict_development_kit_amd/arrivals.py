"""Sensor arrival-time histograms on the GPU (mlgpu_histogram_arrivals).

merlict assigns a photon to the sensor of its final frame whatever its final
interaction (merlict/sensor/Sensor.cpp:14-30) and records its arrival time;
here the arrivals of a batch are binned per sensor and time on the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import torch

from . import _native as N
from .propagation import PropagationResult, _stream_ptr


@dataclass(frozen=True)
class ArrivalBinning:
    """num_bins bins of bin_width seconds from t_start, per sensor; an
    optional arrival-time jitter N(0, jitter_sigma) from Philox stream 2 of
    (seed, photon_id_offset + i)."""
    num_sensors: int
    num_bins: int
    t_start: float
    bin_width: float
    jitter_sigma: float = 0.0
    seed: int = 0
    photon_id_offset: int = 0

    def c_struct(self, counts: torch.Tensor) -> N.Histogram:
        return N.Histogram(counts.data_ptr(), self.num_sensors, self.num_bins, self.t_start, self.bin_width,
                           self.jitter_sigma, self.seed, self.photon_id_offset)

    def empty(self, device) -> torch.Tensor:
        """Zeroed counts (num_sensors, num_bins), uint32 stored as int32."""
        return torch.zeros((self.num_sensors, self.num_bins), dtype=torch.int32, device=device)


def histogram_arrivals(result: PropagationResult, binning: ArrivalBinning, counts: Optional[torch.Tensor] = None,
                       stream=None) -> torch.Tensor:
    """Add the arrivals of `result` to `counts` (allocated zeroed when None)."""
    dev = result.sensor_id.device
    if dev.type != "cuda":
        raise ValueError("results must be on the GPU")
    if counts is None:
        counts = binning.empty(dev)
    if counts.shape != (binning.num_sensors, binning.num_bins) or counts.dtype != torch.int32 \
            or not counts.is_contiguous() or counts.device != dev:
        raise ValueError("counts must be a contiguous int32 (num_sensors, num_bins) tensor on the results' device")
    rs = result.c_struct()
    h = binning.c_struct(counts)
    N.check(N.lib().mlgpu_histogram_arrivals(C.byref(rs), len(result.sensor_id), C.byref(h), _stream_ptr(stream)),
            "mlgpu_histogram_arrivals")
    return counts
