"""The benchmark configurations at their stated sizes, through the HIP path
and against the oracle.

* Lens bench: 1e6 photons of the parallel beam (support uniform on a 0.5 m
  disc 1 m above the lens, tilt uniform in [0, 2] deg, 433 nm), seed 0,
  per photon and as the 19-cell histogram.
* Batched plenoscope bunches: bunches of log-uniform size, each pointed up
  to 3 deg off axis, propagated as one batch by a bunch plan (hipGraph
  replay) and binned per lixel with a 5 ns arrival jitter; the oracle
  propagates and bins the same photons.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
import photon_sets
from merlict_development_kit_amd import (ArrivalBinning, Bunch, GpuScene, PropagationConfig, SceneDescription,
                                         sources)
from merlict_development_kit_amd.bunches import BunchPlan, generate_bunches
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def test_gpu_lens_bench_full_size_matches_oracle(cuda, oracle_threads):
    n = 1_000_000
    desc = SceneDescription.lens()
    ph = photon_sets.lens_beam(n, seed=0)
    gs = GpuScene(desc)
    res = gs.propagate(ph.to(cuda), PropagationConfig(5, seed=0)).numpy()
    ref = O.OracleScene(desc).propagate(ph.numpy(), seed=0, threads=oracle_threads)
    compare(res, ref, n, "lens bench, 1e6 photons")
    cells = np.bincount(res["sensor_id"][res["sensor_id"] >= 0], minlength=19)
    want = np.bincount(ref["sensor_id"][ref["sensor_id"] >= 0], minlength=19)
    print(f"lens bench: {cells.sum()} of {n} photons on the 19 cells")
    assert len(cells) == 19 and cells.sum() > 0.5 * n
    np.testing.assert_array_equal(cells, want)


@pytest.mark.parametrize("schedule", ["wavefront", "persistent"])
def test_gpu_plenoscope_batched_bunches_match_oracle(cuda, schedule, oracle_threads):
    desc = SceneDescription.plenoscope()
    gs = GpuScene(desc)
    gs.set_tuning(schedule=schedule)
    rng = np.random.default_rng(11)
    m = 12
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(2e4), m)).astype(np.int64)
    off = np.radians(3.0) * np.sqrt(rng.uniform(0, 1, m))
    phi = rng.uniform(0, 2 * np.pi, m)
    bunches = [Bunch(sources.plenoscope_bunch(float(off[k] * np.cos(phi[k])), float(off[k] * np.sin(phi[k]))),
                     seed=500 + k, num_photons=int(sizes[k])) for k in range(m)]
    total = int(sizes.sum())
    cfg = PropagationConfig(10, seed=21, photon_id_offset=3000)
    probe = gs.propagate(sources.plenoscope_bunch().generate(20_000, seed=1, device=cuda), PropagationConfig(10))
    t0 = float(probe.time_of_flight[probe.sensor_id >= 0].mean())
    binning = ArrivalBinning(515_023, 64, t0 - 32e-9, 1e-9, jitter_sigma=5e-9, seed=3, photon_id_offset=3000)
    plan = BunchPlan(gs, m, total, cfg, binning, device=cuda)
    plan.counts.zero_()
    assert plan.run(bunches, cfg) == total
    torch.cuda.synchronize()
    got = plan.results(total).numpy()
    got.pop("accumulated_distance", None)  # a plan's results hold no distances
    print(f"plenoscope batch ({schedule}): {m} bunches, {total} photons, "
          f"{int((got['sensor_id'] >= 0).sum())} on lixels", flush=True)
    photons = generate_bunches(bunches, device=cuda).numpy()
    ref = O.OracleScene(desc).propagate(photons, seed=21, max_num_interactions=10, photon_id_offset=3000,
                                        threads=oracle_threads)
    compare(got, ref, total, f"plenoscope batch ({schedule}) vs oracle")
    want = O.histogram_arrivals(ref, 515_023, 64, t0 - 32e-9, 1e-9, 5e-9, seed=3, photon_id_offset=3000)
    counts = plan.counts.cpu().numpy().view(np.uint32)
    print(f"plenoscope batch ({schedule}): {int(counts.sum())} binned, {int(want.sum())} by the oracle")
    assert want.sum() > 0
    np.testing.assert_array_equal(counts, want)
