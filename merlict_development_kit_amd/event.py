"""One plenoscope event on the GPU, step for step as merlict's plenoscope
propagation app runs it (merlict_portal_plenoscope/apps/
plenoscope_propagation.cpp:190-260): CORSIKA photon bunches to photons,
propagation, per-lixel photon pipelines, night-sky background, the
photo-electric converter, single-photon extraction and the photon stream."""
from __future__ import annotations

import json
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import corsika, signal_chain as SC
from .propagation import GpuScene, PropagationConfig


@dataclass
class EventCard:
    """The app's configuration card (the JSON of the plenoscope propagation
    example, plenoscope_propagation.cpp:139-160)."""
    flux_x: np.ndarray
    flux_y: np.ndarray
    nsb_exposure_time: float
    qe_x: np.ndarray
    qe_y: np.ndarray
    dark_rate: float
    crosstalk: float
    time_slice_duration: float
    arrival_time_std: float

    def with_padded_quantum_efficiency(self, lo: float = 200e-9, hi: float = 1200e-9) -> "EventCard":
        """A copy whose quantum efficiency table is extended by zero
        efficiency at lo and hi: the example card tabulates it from 266 nm
        while its night-sky flux starts at 240 nm, so the reference's
        converter throws for background photons below 266 nm."""
        qx = np.concatenate([[lo], self.qe_x, [hi]])
        qy = np.concatenate([[0.0], self.qe_y, [0.0]])
        return EventCard(self.flux_x, self.flux_y, self.nsb_exposure_time, qx, qy, self.dark_rate, self.crosstalk,
                         self.time_slice_duration, self.arrival_time_std)

    @classmethod
    def from_json(cls, path) -> "EventCard":
        card = json.loads(Path(path).read_text())
        nsb, pec, phs = card["night_sky_background_ligth"], card["photo_electric_converter"], card["photon_stream"]
        flux = np.asarray(nsb["flux_vs_wavelength"], np.float64)
        qe = np.asarray(pec["quantum_efficiency_vs_wavelength"], np.float64)
        return cls(flux[:, 0].copy(), flux[:, 1].copy(), float(nsb["exposure_time"]), qe[:, 0].copy(),
                   qe[:, 1].copy(), float(pec["dark_rate"]), float(pec["probability_for_second_puls"]),
                   float(phs["time_slice_duration"]), float(phs["single_photon_arrival_time_resolution"]))


@dataclass
class EventResult:
    symbols: torch.Tensor        # uint8 photon-stream symbols (PhotonStream::write's body)
    truth: torch.Tensor          # int32 truth id per kept pulse (write_simulation_truth)
    nsb_exposure_start_time: float
    num_photons: int             # Cherenkov photons made from the bunches
    num_detected: int            # of those, on a lixel
    num_pulses: int


def propagate_event(scene: GpuScene, bunches: torch.Tensor, num_lixel: int, efficiencies: torch.Tensor,
                    nsb_rate: float, card: EventCard, seed: int, config: PropagationConfig = None) -> EventResult:
    """One event (plenoscope_propagation.cpp:190-260).  bunches: (n, 8)
    float32 CORSIKA photon bunches on the GPU; efficiencies: float32 per
    lixel (calibration.read_efficiencies); nsb_rate: the night-sky light's
    rate over the sensor (signal_chain.nsb_light_rate).  Every step draws
    from its own Philox stream of `seed`, so an event is a function of its
    inputs and seed."""
    cfg = config or PropagationConfig(5, seed=seed)
    photons, truth = corsika.photons_from_bunches(bunches, seed=seed)
    result = scene.propagate(photons, PropagationConfig(cfg.max_num_interactions_per_photon, seed, 0))
    pipes = SC.sensor_pipelines(result.sensor_id, result.error, result.time_of_flight, photons.wavelength, num_lixel,
                                truth_id=truth)
    lit, start = SC.inject_nsb(pipes, num_lixel, efficiencies, card.flux_x, card.flux_y, nsb_rate,
                               card.nsb_exposure_time, seed=seed)
    pulses = SC.photoelectric_convert(lit, num_lixel, card.qe_x, card.qe_y, card.dark_rate, card.crosstalk,
                                      card.nsb_exposure_time, seed=seed)
    slices = SC.extract_pulses(pulses, num_lixel, card.time_slice_duration, card.arrival_time_std, seed=seed)
    symbols, kept_truth = SC.photon_stream(pulses, slices, num_lixel)
    return EventResult(symbols, kept_truth, start, len(photons.wavelength), len(pipes.arrival_time),
                       len(pulses.arrival_time))


def plenoscope_nsb_rate(card: EventCard, max_fov_diameter_deg: float = 6.5,
                        max_aperture_radius: float = 35.5) -> float:
    """The night-sky rate over the plenoscope's light-field sensor
    (night_sky_background::Light of its geometry)."""
    return SC.nsb_light_rate(card.flux_x, card.flux_y, math.radians(max_fov_diameter_deg) / 2.0,
                             max_aperture_radius)
