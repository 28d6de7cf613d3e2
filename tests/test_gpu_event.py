"""One plenoscope event through every step on the GPU (event.propagate_event:
CORSIKA bunches, propagation, pipelines, night-sky background, converter,
extraction, photon stream) against the same chain of oracle steps."""
import math

import numpy as np
import pytest
import torch

import pyoracle as O
from conftest import GOLDEN
from merlict_development_kit_amd import GpuScene, SceneDescription
from merlict_development_kit_amd import event as EV
from merlict_development_kit_amd import signal_chain as SC
from merlict_development_kit_amd.sources import Source

pytestmark = pytest.mark.gpu

LIXELS = 515_023  # lixels of the default plenoscope's light-field sensor


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def synthetic_bunches(n, seed):
    """CORSIKA-like bunches over the plenoscope's aperture: x, y within 30 m
    (cm), direction cosines within a few mrad, 0-5 ns, 10 km emission
    height, weights 0.5-3, wavelengths 300-550 nm (negative, as CORSIKA
    writes them)."""
    rng = np.random.default_rng(seed)
    r, phi = 3000.0 * np.sqrt(rng.uniform(0, 1, n)), rng.uniform(0, 2 * np.pi, n)
    b = np.empty((n, 8), np.float32)
    b[:, 0], b[:, 1] = r * np.cos(phi), r * np.sin(phi)
    b[:, 2], b[:, 3] = rng.normal(0, 0.004, n), rng.normal(0, 0.004, n)
    b[:, 4] = rng.uniform(0, 5, n)
    b[:, 5] = 1e6
    b[:, 6] = rng.uniform(0.5, 3.0, n)
    b[:, 7] = -rng.uniform(300, 550, n)
    return b


def oracle_event(desc, bunches, eff, rate, card, seed):
    ph, truth = O.corsika_photons(bunches, seed=seed)
    res = O.OracleScene(desc).propagate(ph, 5, seed, 0, threads=8)
    off, t, lam, tr = O.sensor_pipelines(res["sensor_id"], res["error"], res["time_of_flight"], ph["wavelength"],
                                         truth, LIXELS)
    off, t, lam, tr, start = O.inject_nsb(off, t, lam, tr, eff, card.flux_x, card.flux_y, rate,
                                          card.nsb_exposure_time, seed)
    po, pt, ptr = O.photoelectric(off, t, lam, tr, LIXELS, card.qe_x, card.qe_y, card.dark_rate, card.crosstalk,
                                  card.nsb_exposure_time, seed)
    sl = O.extract_pulses(po, pt, LIXELS, card.time_slice_duration, card.arrival_time_std, seed)
    kept = sl >= 0
    # channel by channel the kept slices, a marker (255) at the end of every
    # channel but the last
    sensor = np.searchsorted(po, np.arange(len(sl)), side="right") - 1
    per = np.bincount(sensor[kept], minlength=LIXELS)
    ends = np.cumsum(per)[:-1]
    sym = np.insert(sl[kept].astype(np.uint8), ends, np.uint8(255))
    return sym, ptr[kept], start


def test_gpu_plenoscope_event_matches_oracle_chain(cuda):
    """4,000 bunches (about 7,000 photons) on the plenoscope with a weak
    night sky: the photon stream and its truth ids equal those of the
    oracle's chain of the same steps, and a second run gives the same
    event."""
    desc = SceneDescription.plenoscope()
    gs = GpuScene(desc)
    card = EV.EventCard.from_json(GOLDEN / "plenoscope_propagation_input_card.json").with_padded_quantum_efficiency()
    bunches = synthetic_bunches(4000, seed=3)
    eff = np.random.default_rng(4).uniform(0.5, 1.0, LIXELS).astype(np.float32)
    rate = 2e11  # about 0.02 background photons per lixel
    b = torch.tensor(bunches, device=cuda)
    e = torch.tensor(eff, device=cuda)
    ev = EV.propagate_event(gs, b, LIXELS, e, rate, card, seed=9)
    again = EV.propagate_event(gs, b, LIXELS, e, rate, card, seed=9)
    assert torch.equal(ev.symbols, again.symbols) and torch.equal(ev.truth, again.truth)
    assert ev.num_detected > 1000
    sym, truth, start = oracle_event(desc, bunches, eff, rate, card, 9)
    assert ev.nsb_exposure_start_time == start
    np.testing.assert_array_equal(ev.symbols.cpu().numpy(), sym)
    np.testing.assert_array_equal(ev.truth.cpu().numpy(), truth)
    t = ev.truth.cpu().numpy()
    assert ((t >= 0) & (t < 4000) | np.isin(t, [-100, -201, -202])).all()
    assert (t >= 0).sum() > 100 and (t == -100).sum() > 0
