"""Plenoscope light-field calibration on the GPU
(merlict_portal_plenoscope/calibration): calibration photons from the
aperture, propagated, reduced to per-lixel statistics, and the
lixel_statistics.bin file format (include/mlgpu.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np
import torch

from . import _native as N
from . import sources
from .propagation import GpuScene, PropagationConfig, PropagationResult

# plenoscope::CalibrationPhotonResult (calibration/CalibrationPhotonResult.h:9-19)
CALIBRATION_PHOTON = np.dtype([("reached_sensor", "<u4"), ("lixel_id", "<u4"),
                               ("x_pos_on_principal_aperture", "<f4"), ("y_pos_on_principal_aperture", "<f4"),
                               ("x_tilt_vs_optical_axis", "<f4"), ("y_tilt_vs_optical_axis", "<f4"),
                               ("time_of_flight", "<f4")])
# plenoscope::calibration::LixelStatistic (calibration/LixelStatistics.h:16-24)
LIXEL_STATISTIC = np.dtype([(k, "<f4") for k in (
    "efficiency", "efficiency_std", "cx_mean", "cx_std", "cy_mean", "cy_std", "x_mean", "x_std", "y_mean", "y_std",
    "time_delay_mean", "time_delay_std")])

# night_sky_background/NightSkyBackground.h:11-12 and Calibrator.h:27-28
APERTURE_RADIUS_OVERHEAD = 1.05
FOV_RADIUS_OVERHEAD = 1.20
DISTANCE_TO_APERTURE_PLANE = 1e4
WAVELENGTH = 433e-9


@dataclass
class CalibrationConfig:
    """calibration::Config (calibration/Config.cpp:6-9)."""
    num_blocks: int = 8
    photons_per_block: int = 1_000_000


def calibration_source(max_aperture_radius: float, max_fov_radius: float) -> sources.Source:
    """The photons of Calibrator (Calibrator.cpp:19-60): aperture point
    uniform on a disc of 1.05 x the aperture radius, incident direction
    uniform on the cap within 1.2 x the FoV radius, started 1e4 m upstream,
    433 nm."""
    return sources.calibration(APERTURE_RADIUS_OVERHEAD * max_aperture_radius, FOV_RADIUS_OVERHEAD * max_fov_radius,
                               DISTANCE_TO_APERTURE_PLANE, WAVELENGTH)


def _st(stream, dev):
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    return C.c_void_p(s.cuda_stream)


def calibration_photons(source: sources.Source, seed: int, photon_id_offset: int, result: PropagationResult,
                        num_lixel: int, stream=None) -> torch.Tensor:
    """Records (n x 28 bytes, as int32 words on the GPU) of one_photon
    (Calibrator.cpp:48-102) for the photons photon_id_offset .. of `source`
    and their results (mlgpu_calibration_photons); numpy view:
    ``records.cpu().numpy().view(CALIBRATION_PHOTON)``."""
    n = len(result.sensor_id)
    dev = result.sensor_id.device
    out = torch.empty(n * 7, dtype=torch.int32, device=dev)
    rs = result.c_struct()
    N.check(N.lib().mlgpu_calibration_photons(C.byref(source.to_c()), seed, photon_id_offset, C.byref(rs), n,
                                              num_lixel, C.c_void_p(out.data_ptr() if n else None), _st(stream, dev)),
            "mlgpu_calibration_photons")
    return out


class LixelStatisticsFiller:
    """LixelStatisticsFiller (LixelStatisticsFiller.cpp:8-62) with its
    online state on the GPU (mlgpu_lixel_filler_*)."""

    def __init__(self, num_lixel: int, num_blocks: int, num_photons_per_block: int):
        h = C.c_void_p()
        N.check(N.lib().mlgpu_lixel_filler_create(num_lixel, num_blocks, num_photons_per_block, C.byref(h)),
                "mlgpu_lixel_filler_create")
        self._h = h
        self.num_lixel = num_lixel

    def fill_in_block(self, records: torch.Tensor, stream=None) -> None:
        """Adds a block of records (int32 words, 7 per record, on the GPU)."""
        if records.dtype != torch.int32 or records.device.type != "cuda" or records.numel() % 7:
            raise ValueError("records must be int32 words on the GPU, 7 per record")
        r = records.contiguous()
        N.check(N.lib().mlgpu_lixel_filler_fill_block(self._h, C.c_void_p(r.data_ptr() if r.numel() else None),
                                                      r.numel() // 7, _st(stream, r.device)),
                "mlgpu_lixel_filler_fill_block")

    def get_lixel_statistics(self, device="cuda", stream=None) -> np.ndarray:
        """The num_lixel LixelStatistic records (numpy, LIXEL_STATISTIC)."""
        dev = torch.device(device)
        out = torch.empty(self.num_lixel * 12, dtype=torch.float32, device=dev)
        N.check(N.lib().mlgpu_lixel_filler_statistics(self._h, C.c_void_p(out.data_ptr()), _st(stream, dev)),
                "mlgpu_lixel_filler_statistics")
        return out.cpu().numpy().view(LIXEL_STATISTIC)

    def close(self) -> None:
        if self._h:
            N.lib().mlgpu_lixel_filler_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_calibration(scene: GpuScene, num_lixel: int, max_aperture_radius: float, max_fov_radius: float,
                    config: CalibrationConfig = CalibrationConfig(), seed: int = 0,
                    propagation: Optional[PropagationConfig] = None, world_frame: int = 0,
                    device="cuda") -> np.ndarray:
    """run_calibration (Calibrator.cpp:140-163): num_blocks blocks of
    photons_per_block calibration photons, propagated (5 interactions, the
    reference's default) and added to the filler block by block; the
    per-lixel statistics.  Photon b * photons_per_block + i of the run draws
    from Philox (seed, id) for its source and its propagation."""
    src = calibration_source(max_aperture_radius, max_fov_radius)
    cfg = propagation or PropagationConfig(5, seed=seed)
    filler = LixelStatisticsFiller(num_lixel, config.num_blocks, config.photons_per_block)
    try:
        n = config.photons_per_block
        for b in range(config.num_blocks):
            offset = b * n
            photons = src.generate(n, seed=seed, photon_id_offset=offset, device=device)
            res = scene.propagate(photons, PropagationConfig(cfg.max_num_interactions_per_photon, cfg.seed, offset),
                                  world_frame=world_frame)
            filler.fill_in_block(calibration_photons(src, seed, offset, res, num_lixel))
        return filler.get_lixel_statistics(device)
    finally:
        filler.close()


def write(lixel_statistics: np.ndarray, path) -> None:
    """calibration::write (LixelStatistics.cpp:18-32): the records, raw."""
    Path(path).write_bytes(np.ascontiguousarray(lixel_statistics, LIXEL_STATISTIC).tobytes())


def read(path) -> np.ndarray:
    """calibration::read (LixelStatistics.cpp:34-55): whole records only."""
    raw = Path(path).read_bytes()
    k = len(raw) // LIXEL_STATISTIC.itemsize
    return np.frombuffer(raw[:k * LIXEL_STATISTIC.itemsize], LIXEL_STATISTIC).copy()


def read_efficiencies(path) -> np.ndarray:
    """calibration::read_efficiencies (LixelStatistics.cpp:57-80)."""
    return read(path)["efficiency"].astype(np.float32)
