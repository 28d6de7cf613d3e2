"""GPU photon sources (mlgpu_generate_photons) against the oracle's
restatement: every photon bit for bit, since both use only correctly rounded
IEEE operations on the same Philox draws."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import pyoracle as O
from merlict_development_kit_amd import sources as S
from merlict_development_kit_amd._native import MlgpuError

pytestmark = pytest.mark.gpu

SOURCES = {
    "lens_beam": S.lens_beam(),
    "fact_bunch": S.fact_bunch(),
    "calibration": S.calibration(),
    "point_like": S.point_like(math.radians(10.0)),
    "parallel_disc": S.parallel_disc(0.3),
    "posed_bunch": S.fact_bunch().posed(S.rotation_xyz(0.02, -0.03, 1.0), (0.5, -0.25, 0.0)),
    "shower_ring": S.shower_ring(0.9),
    "night_sky": S.night_sky(),
}


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.mark.parametrize("name", sorted(SOURCES))
def test_gpu_source_matches_oracle_bitwise(cuda, name):
    src = SOURCES[name]
    n = 100_000
    got = src.generate(n, seed=21, photon_id_offset=5, device=cuda).numpy()
    want = O.generate_photons(src.to_c(O.Source), n, seed=21, photon_id_offset=5)
    for k in want:
        mism = int((got[k] != want[k]).sum())
        assert mism == 0, f"{name}: {mism} of {n} photons differ in {k}"


def test_gpu_source_rejects_bad_parameters(cuda):
    for bad in (S.Source(direction="tilt", angle=2.0), S.Source(lambda_min=0.0, lambda_max=1e-7),
                S.Source(lambda_min=500e-9, lambda_max=400e-9), S.Source(support="disc", radius=-1.0)):
        with pytest.raises(MlgpuError):
            bad.generate(10, device=cuda)
    assert len(S.fact_bunch().generate(0, device=cuda)) == 0
    with pytest.raises(ValueError):
        S.fact_bunch().generate(10, device="cpu")


def test_gpu_arrival_histograms_match_oracle(cuda):
    """Arrival histograms of a propagated FACT bunch: the GPU counts equal
    the oracle's binning of the same results, with and without jitter, for
    the LDS (1440 pixels x 8 bins) and the global-atomic (x 64 bins) paths."""
    from merlict_development_kit_amd import (ArrivalBinning, GpuScene, PropagationConfig, SceneDescription,
                                             histogram_arrivals)
    gs = GpuScene(SceneDescription.fact(light_guides=True))
    n = 500_000
    res = gs.propagate(S.fact_bunch().generate(n, seed=3, device=cuda), PropagationConfig(5, seed=4))
    host = res.numpy()
    for bins, jitter in ((8, 0.0), (8, 2e-9), (64, 0.0), (64, 5e-9)):
        b = ArrivalBinning(1440, bins, 40e-9, 32e-9 / bins, jitter, seed=9, photon_id_offset=100)
        got = histogram_arrivals(res, b).cpu().numpy().view(np.uint32)
        want = O.histogram_arrivals(host, 1440, bins, 40e-9, 32e-9 / bins, jitter, seed=9, photon_id_offset=100)
        assert want.sum() > 0.25 * n
        np.testing.assert_array_equal(got, want, err_msg=f"{bins} bins, jitter {jitter}")
    # counts accumulate
    b = ArrivalBinning(1440, 8, 48e-9, 2e-9)
    twice = histogram_arrivals(res, b, histogram_arrivals(res, b))
    np.testing.assert_array_equal(twice.cpu().numpy().view(np.uint32), 2 * O.histogram_arrivals(host, 1440, 8, 48e-9, 2e-9))


def test_gpu_bunches_equal_separate_draws_and_oracle(cuda, oracle_threads):
    """mlgpu_propagate_bunches: the same photons as drawing each bunch on its
    own, the same results as propagating their concatenation (and the
    oracle's), the same histogram as binning those results."""
    from merlict_development_kit_amd import (ArrivalBinning, Bunch, GpuScene, PhotonBatch, PropagationConfig,
                                             SceneDescription, propagate_bunches)
    from merlict_development_kit_amd.bunches import generate_bunches
    desc = SceneDescription.fact(light_guides=True)
    gs = GpuScene(desc)
    bunches = [Bunch(S.fact_bunch().posed(S.rotation_xyz(0.01 * k, -0.02 * k, 0.0)), seed=40 + k, num_photons=m)
               for k, m in enumerate((30_000, 0, 1, 77_777, 12_345))]
    cfg = PropagationConfig(5, seed=8, photon_id_offset=1000)
    binning = ArrivalBinning(1440, 16, 48e-9, 1e-9, jitter_sigma=1e-9, seed=2, photon_id_offset=5)
    for schedule in ("fused", "wavefront", "persistent"):
        gs.set_tuning(schedule=schedule)
        counts, res = propagate_bunches(gs, bunches, cfg, binning, want_results=True, device=cuda)
        parts = [b.source.generate(b.num_photons, seed=b.seed, device=cuda) for b in bunches]
        cat = PhotonBatch(*[torch.cat([getattr(p, f) for p in parts]) for f in
                            ("support_x", "support_y", "support_z", "direction_x", "direction_y", "direction_z",
                             "wavelength")])
        again = gs.propagate(cat, cfg).numpy()
        gen = generate_bunches(bunches, device=cuda).numpy()
        for f in O.PHOTON_FIELDS:
            np.testing.assert_array_equal(gen[f], getattr(cat, f).cpu().numpy(), err_msg=f"generate_bunches {f}")
        got = res.numpy()
        for k in got:
            np.testing.assert_array_equal(got[k], again[k], err_msg=f"{schedule} {k}")
        want = O.histogram_arrivals(got, 1440, 16, 48e-9, 1e-9, 1e-9, seed=2, photon_id_offset=5)
        np.testing.assert_array_equal(counts.cpu().numpy().view(np.uint32), want)
    host = {k: np.concatenate([O.generate_photons(b.source.to_c(O.Source), b.num_photons, seed=b.seed)[k]
                               for b in bunches]) for k in O.PHOTON_FIELDS}
    ref = O.OracleScene(desc).propagate(host, seed=8, photon_id_offset=1000, threads=oracle_threads)
    for k in ("final_frame", "sensor_id", "final_interaction", "num_interactions"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    # histogram only, no per-photon results
    counts2, none = propagate_bunches(gs, bunches, cfg, binning, device=cuda)
    assert none is None
    np.testing.assert_array_equal(counts2.cpu().numpy(), counts.cpu().numpy())


@pytest.mark.parametrize("schedule", ["wavefront", "persistent"])
def test_gpu_bunch_plan_replays_match_eager(cuda, schedule):
    """A bunch plan (one hipGraph, replayed per batch with the batch's size,
    seed and id offset read from device memory) gives, replay after replay,
    the histograms and per-photon results of mlgpu_propagate_bunches on the
    same batch; other work runs eagerly between the replays.  The graph holds
    the wavefront kernels or the one persistent work-queue kernel."""
    from merlict_development_kit_amd import (ArrivalBinning, Bunch, GpuScene, PropagationConfig, SceneDescription,
                                             propagate_bunches)
    from merlict_development_kit_amd._native import MlgpuError
    from merlict_development_kit_amd.bunches import BunchPlan
    gs = GpuScene(SceneDescription.fact(light_guides=True))
    gs.set_tuning(schedule=schedule)
    binning = ArrivalBinning(1440, 16, 48e-9, 1e-9, jitter_sigma=1e-9, seed=2, photon_id_offset=5)
    plan = BunchPlan(gs, max_bunches=8, max_photons=300_000, config=PropagationConfig(5, seed=1), binning=binning,
                     device=cuda)
    batches = [
        [Bunch(S.fact_bunch().posed(S.rotation_xyz(0.01, 0.0, 0.0)), seed=3, num_photons=120_000),
         Bunch(S.shower_ring(0.8), seed=4, num_photons=50_001)],
        [Bunch(S.night_sky(), seed=5, num_photons=7)],
        [Bunch(S.fact_bunch(), seed=6 + k, num_photons=37_000 + 1000 * k) for k in range(7)],
    ]
    for rep in range(2):
        for j, batch in enumerate(batches):
            cfg = PropagationConfig(5, seed=100 + j, photon_id_offset=1000 * j)
            want, res = propagate_bunches(gs, batch, cfg, binning, want_results=True, device=cuda)
            torch.cuda.synchronize()
            plan.counts.zero_()
            n = plan.run(batch, cfg)
            torch.cuda.synchronize()
            assert torch.equal(plan.counts, want), (rep, j)
            got = plan.results(n).numpy()
            ref = res.numpy()
            for k in got:
                if k != "accumulated_distance":
                    np.testing.assert_array_equal(got[k], ref[k], err_msg=f"replay {rep}, batch {j}, {k}")
            # eager work between replays
            gs.propagate(S.fact_bunch().generate(20_000, seed=j, device=cuda), PropagationConfig(5, seed=j))
    with pytest.raises(MlgpuError):
        plan.run(batches[2] + batches[0], PropagationConfig(5))
    with pytest.raises(MlgpuError):
        plan.run(batches[1], PropagationConfig(4))
    with pytest.raises(MlgpuError):
        plan.run([Bunch(S.fact_bunch(), seed=1, num_photons=300_001)], PropagationConfig(5))
