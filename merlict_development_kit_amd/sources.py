"""Photon sources drawn on the GPU (mlgpu_generate_photons).

Every photon comes from Philox4x32-10 stream 1 of (seed, photon id), so a
bunch is reproducible from its seed and id range alone, on any GPU and by
the CPU restatement.  The presets are the synthetic inputs of the measured
configurations and the reference's own sources:

* :func:`lens_beam` -- supports on a disc above a lens, tilt U[0, max],
  azimuth U[0, 2 pi), one wavelength;
* :func:`fact_bunch` -- supports on a 4.5 m disc above the dish, x and y
  tilts N(0, 0.5 deg), lambda in [280, 700] nm with pdf ~ lambda^-2;
* :func:`calibration` -- plenoscope calibration photons
  (merlict_portal_plenoscope/calibration/Calibrator.cpp:37-102);
* :func:`plenoscope_bunch` -- a photon bunch for the plenoscope with a
  per-bunch pointing offset;
* :func:`point_like` and :func:`parallel_disc` --
  photon_source::point_like_towards_z_opening_angle_num_photons and
  parallel_towards_z_from_xy_disc (merlict/Photons.cpp:85-125).

A pose (rotation, position) applied to supports and directions restates
transform_photons (merlict_json/json.cpp:512-519).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, replace
from typing import Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from .propagation import PhotonBatch

SUPPORTS = {"point": 0, "disc": 1}
DIRECTIONS = {"axis": 0, "tilt": 1, "gaussian": 2, "cap": 3, "ring": 4}
SPECTRA = {"inverse_square": 0, "uniform": 1}

IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
# rotation by pi about x: directions about +z become directions about -z
FLIP = (1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, -1.0)


def rotation_xyz(rx: float, ry: float, rz: float) -> Tuple[float, ...]:
    """Row-major rotation of HomTra3 from xyz angles, the T that
    set_rotation_component_based_on_xyz_angles fills (merlict/HomTra3.cpp:93-112)."""
    cx, sx, cy, sy, cz, sz = math.cos(rx), math.sin(rx), math.cos(ry), math.sin(ry), math.cos(rz), math.sin(rz)
    return (cy * cz, cx * sz + sx * sy * cz, sx * sz - cx * sy * cz,
            -cy * sz, cx * cz - sx * sy * sz, sx * cz + cx * sy * sz,
            sy, -sx * cy, cx * cy)


def compose(a: Sequence[float], b: Sequence[float]) -> Tuple[float, ...]:
    """Row-major product a b of two 3x3 rotations."""
    A = np.asarray(a, dtype=np.float64).reshape(3, 3)
    B = np.asarray(b, dtype=np.float64).reshape(3, 3)
    return tuple((A @ B).reshape(9).tolist())


@dataclass(frozen=True)
class Source:
    """mlgpu_source: support on the xy plane, direction about +z, wavelength,
    start distance back along the direction, then the pose."""
    support: str = "point"
    direction: str = "axis"
    radius: float = 0.0
    angle: float = 0.0
    lambda_min: float = 433e-9
    lambda_max: float = 433e-9
    start_distance: float = 0.0
    rot: Tuple[float, ...] = IDENTITY
    pos: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    angle_width: float = 0.0
    spectrum: str = "inverse_square"

    def to_c(self, struct_type=N.Source):
        """The fields as a ctypes struct laid out as mlgpu_source."""
        s = struct_type()
        s.support = SUPPORTS[self.support]
        s.direction = DIRECTIONS[self.direction]
        s.radius = self.radius
        s.angle = self.angle
        s.lambda_min = self.lambda_min
        s.lambda_max = self.lambda_max
        s.start_distance = self.start_distance
        for k in range(9):
            s.rot[k] = self.rot[k]
        for k in range(3):
            s.pos[k] = self.pos[k]
        s.angle_width = self.angle_width
        s.spectrum = SPECTRA[self.spectrum]
        return s

    def posed(self, rot: Sequence[float] = IDENTITY, pos: Sequence[float] = (0.0, 0.0, 0.0)) -> "Source":
        """The source moved by a further pose (applied after its own)."""
        R = np.asarray(rot, dtype=np.float64).reshape(3, 3)
        new_pos = R @ np.asarray(self.pos, dtype=np.float64) + np.asarray(pos, dtype=np.float64)
        return replace(self, rot=compose(rot, self.rot), pos=tuple(new_pos.tolist()))

    def generate(self, n: int, seed: int = 0, photon_id_offset: int = 0, device="cuda",
                 stream=None) -> PhotonBatch:
        """Photons photon_id_offset .. + n - 1, drawn on the GPU."""
        device = torch.device(device)
        if device.type != "cuda":
            raise ValueError("photon sources draw on the GPU; copy the batch to the host afterwards")
        arrays = [torch.empty(n, dtype=torch.float64, device=device) for _ in range(7)]
        if n > 0:
            buf = N.PhotonBuffers(*[a.data_ptr() for a in arrays])
            st = stream if stream is not None else torch.cuda.current_stream(device)
            N.check(N.lib().mlgpu_generate_photons(C.byref(self.to_c()), n, seed, photon_id_offset, C.byref(buf),
                                                   C.c_void_p(st.cuda_stream)), "mlgpu_generate_photons")
        return PhotonBatch(*arrays)


def lens_beam(disc_radius: float = 0.25, height: float = 1.0, max_tilt_deg: float = 2.0,
              wavelength: float = 433e-9) -> Source:
    """Supports on a disc at z = height, towards -z tilted by U[0, max_tilt]."""
    return Source("disc", "tilt", disc_radius, math.radians(max_tilt_deg), wavelength, wavelength, 0.0,
                  FLIP, (0.0, 0.0, height))


def fact_bunch(disc_radius: float = 2.25, height: float = 10.0, sigma_deg: float = 0.5,
               lambda_min: float = 280e-9, lambda_max: float = 700e-9) -> Source:
    """Supports on a disc at z = height, towards -z with x and y tilts
    N(0, sigma), lambda with pdf ~ lambda^-2."""
    return Source("disc", "gaussian", disc_radius, math.radians(sigma_deg), lambda_min, lambda_max, 0.0,
                  FLIP, (0.0, 0.0, height))


def calibration(aperture_radius: float = 35.5 * 1.05, max_incident_angle: float = math.radians(3.25) * 1.2,
                distance: float = 1e4, wavelength: float = 433e-9) -> Source:
    """Calibration photons: aperture point uniform on a disc, incident
    direction c uniform on the cap within max_incident_angle of +z, started
    `distance` upstream (at aperture point + c distance), travelling -c."""
    return Source("disc", "cap", aperture_radius, max_incident_angle, wavelength, wavelength, distance,
                  FLIP, (0.0, 0.0, 0.0))


def plenoscope_bunch(offset_x: float = 0.0, offset_y: float = 0.0, aperture_radius: float = 35.5 * 1.05,
                     spread_deg: float = 0.05, distance: float = 1e3, lambda_min: float = 280e-9,
                     lambda_max: float = 700e-9) -> Source:
    """One bunch of Cherenkov-like photons for the plenoscope: supports on the
    aperture disc, x and y tilts N(0, spread) about the bunch's arrival
    direction, which is the optical axis turned by the pointing offsets
    (rad, about x then y), started `distance` upstream, lambda with pdf ~
    lambda^-2."""
    src = Source("disc", "gaussian", aperture_radius, math.radians(spread_deg), lambda_min, lambda_max, distance,
                 FLIP, (0.0, 0.0, 0.0))
    return src.posed(rotation_xyz(offset_x, offset_y, 0.0))


def point_like(opening_angle: float, wavelength: float = 433e-9) -> Source:
    """photon_source::point_like_towards_z_opening_angle_num_photons."""
    return Source("point", "cap", 0.0, opening_angle, wavelength, wavelength)


def parallel_disc(disc_radius: float, wavelength: float = 433e-9) -> Source:
    """photon_source::parallel_towards_z_from_xy_disc."""
    return Source("disc", "axis", disc_radius, 0.0, wavelength, wavelength)


def shower_ring(ring_radius_deg: float, width_deg: float = 0.1, disc_radius: float = 2.25, height: float = 10.0,
                lambda_min: float = 280e-9, lambda_max: float = 700e-9) -> Source:
    """A shower-like bunch for a telescope: supports on a disc at z =
    height, directions on a Cherenkov ring about -z (zenith ~ N(radius,
    width), azimuth uniform), lambda with pdf ~ lambda^-2."""
    return Source("disc", "ring", disc_radius, math.radians(ring_radius_deg), lambda_min, lambda_max, 0.0,
                  FLIP, (0.0, 0.0, height), angle_width=math.radians(width_deg))


def night_sky(field_radius_deg: float = 2.0, disc_radius: float = 2.25, height: float = 10.0,
              lambda_min: float = 300e-9, lambda_max: float = 600e-9) -> Source:
    """Night-sky background for a telescope: supports on a disc at z =
    height, directions uniform on the cap within field_radius of -z, lambda
    uniform."""
    return Source("disc", "cap", disc_radius, math.radians(field_radius_deg), lambda_min, lambda_max, 0.0,
                  FLIP, (0.0, 0.0, height), spectrum="uniform")
