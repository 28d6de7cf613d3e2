"""The surrogate + fallback hybrid (mlgpu_propagate_hybrid) on the FACT
telescope with light guides: the fallback path against the plain
propagator (merlict/Photons.cpp:9-22 is the path a fallback photon takes),
the entropy bound against the network's own outputs, and a miniature of the
full-scale stream (Cherenkov-ring bunches plus night-sky background)
against plain propagation of the same photons."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest
import torch

LOG_SIGMA_MIN, LOG_SIGMA_MAX = -9.0, 6.0
HALF_LOG_2PI_E = 1.4189385332046727


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def fact_trained(cuda):
    """A surrogate trained on 1e8 photons of the FACT domain (a few
    seconds on the GPU), with the scene, domain and cells it was trained on."""
    from merlict_development_kit_amd import GpuScene, SceneDescription
    from merlict_development_kit_amd import surrogate as SG
    desc = SceneDescription.fact(light_guides=True)
    gs = GpuScene(desc)
    cells = SG.HexCells(desc, cuda)
    sur = SG.Surrogate(seed=1)
    dom, _ = SG.train(gs, desc, sur, SG.fact_domain(), 100_000_000, seed=3, cells=cells,
                      focal_desc=SceneDescription.fact(focal_plane=True))
    return gs, desc, sur, dom, cells


def _binning(dom, **kw):
    from merlict_development_kit_amd import ArrivalBinning
    return ArrivalBinning(1440, 16, dom.time_offset - 8 * dom.time_scale, dom.time_scale, **kw)


def _entropy_and_domain(sur, dom, photons):
    """Per photon, in float64 from the network's outputs (Surrogate.forward
    on the library's features): the mixture entropy bound the hybrid
    compares with max_entropy (entropy of the weights plus their mean
    component entropy, log sigma clamped to [-9, 6]) and whether the
    features are in the domain."""
    from merlict_development_kit_amd import surrogate as SG
    f, _ = SG.prepare(dom, photons)
    o = sur.forward(f).cpu().numpy().astype(np.float64)
    f = f.cpu().numpy()
    w = o[:, 1::7][:, :8]
    lse = w.max(axis=1, keepdims=True) + np.log(np.exp(w - w.max(axis=1, keepdims=True)).sum(axis=1, keepdims=True))
    p = np.exp(w - lse)
    ls = np.stack([o[:, 5 + 7 * q:8 + 7 * q] for q in range(8)], axis=1)
    hk = 3 * HALF_LOG_2PI_E + np.clip(ls, LOG_SIGMA_MIN, LOG_SIGMA_MAX).sum(axis=2)
    h = (p * (hk - (w - lse))).sum(axis=1)
    inside = ((f[:, 0] * f[:, 0] + f[:, 1] * f[:, 1] <= np.float32(1)) &
              (f[:, 2] * f[:, 2] + f[:, 3] * f[:, 3] <= np.float32(1)) & (f[:, 4] >= -1) & (f[:, 4] <= 1))
    return h, inside


@pytest.mark.gpu
def test_gpu_hybrid_all_fallback_equals_plain(cuda, fact_trained):
    """max_entropy = -inf sends every photon to the propagator: the hybrid's
    histogram then equals, bit for bit, the plain propagation of the same
    batch binned by mlgpu_histogram_arrivals, with arrival-time jitter on and
    nonzero photon id offsets, over a batch of more than one hybrid chunk
    (2^24 photons), so every fallback photon is propagated and jittered
    under its own id."""
    from merlict_development_kit_amd import PropagationConfig, histogram_arrivals, sources
    from merlict_development_kit_amd import surrogate as SG
    gs, desc, sur, dom, cells = fact_trained
    n = (1 << 24) + 300_000
    ph = sources.fact_bunch().generate(n, seed=21, device=cuda)
    cfg = PropagationConfig(5, seed=4, photon_id_offset=12_345)
    binning = _binning(dom, jitter_sigma=0.5e-9, seed=9, photon_id_offset=777)
    plain = histogram_arrivals(gs.propagate(ph, cfg), binning)
    hy, fallback = SG.propagate_hybrid(gs, sur, dom, cells, ph, binning, cfg, max_entropy=float("-inf"), seed=8)
    torch.cuda.synchronize()
    print(f"{n} photons, {fallback} fell back, {int(plain.sum())} binned")
    assert fallback == n
    assert int(plain.sum()) > 0.3 * n
    assert torch.equal(hy, plain)


@pytest.mark.gpu
def test_gpu_hybrid_entropy_bound_selects_the_fallback_photons(cuda, fact_trained):
    """A finite entropy bound at about the 88th percentile of the photons'
    entropies (placed in the widest gap between neighbouring entropies
    there, so no photon lies within float rounding of it): the fallback
    count equals the number of photons above the bound or out of the domain,
    computed on the host from Surrogate.forward; with cells that no draw can
    land on, the hybrid's histogram is exactly the plain histogram of those
    photons; two runs with the real cells are bit-identical."""
    from merlict_development_kit_amd import PropagationConfig, histogram_arrivals, sources
    from merlict_development_kit_amd import surrogate as SG
    gs, desc, sur, dom, cells = fact_trained
    n = 200_000
    # wider than the domain (2.31 m, 2.64 deg), so some photons are outside it
    ph = sources.fact_bunch(disc_radius=2.4, sigma_deg=1.0).generate(n, seed=31, device=cuda)
    h, inside = _entropy_and_domain(sur, dom, ph)
    hs = np.sort(h)
    lo, hi = int(0.85 * n), int(0.92 * n)
    g = int(np.argmax(np.diff(hs[lo:hi]))) + lo
    bound = 0.5 * (hs[g] + hs[g + 1])
    want = (h > bound) | ~inside
    print(f"bound {bound:.5f} in a gap of {hs[g + 1] - hs[g]:.2e}; {int((h > bound).sum())} above it, "
          f"{int((~inside).sum())} out of the domain, {int(want.sum())} to fall back")
    assert hs[g + 1] - hs[g] > 1e-5 * max(1.0, abs(bound))  # float32 entropies are good to ~1e-6
    assert 0.05 * n < want.sum() < 0.35 * n and (~inside).sum() > 0
    cfg = PropagationConfig(5, seed=6, photon_id_offset=1000)
    binning = _binning(dom, jitter_sigma=0.5e-9, seed=2, photon_id_offset=55)
    # no cells: every surrogate draw misses the sensors, only fallbacks count
    nowhere = SG.HexCells.__new__(SG.HexCells)
    nowhere.__dict__.update(cells.__dict__)
    nowhere.table = torch.full_like(cells.table, -1)
    only_fb, k = SG.propagate_hybrid(gs, sur, dom, nowhere, ph, binning, cfg, max_entropy=bound, seed=3)
    assert k == int(want.sum())
    res = gs.propagate(ph, cfg)
    keep = torch.tensor(want, device=cuda)
    res.sensor_id = torch.where(keep, res.sensor_id, torch.full_like(res.sensor_id, -1))
    expect = histogram_arrivals(res, binning)
    torch.cuda.synchronize()
    print(f"fallback photons binned: {int(only_fb.sum())}")
    assert int(expect.sum()) > 0
    assert torch.equal(only_fb, expect)
    # with the real cells: reproducible, and the surrogate adds its draws
    a, ka = SG.propagate_hybrid(gs, sur, dom, cells, ph, binning, cfg, max_entropy=bound, seed=3)
    b, kb = SG.propagate_hybrid(gs, sur, dom, cells, ph, binning, cfg, max_entropy=bound, seed=3)
    assert ka == kb == k
    assert torch.equal(a, b)
    assert int(a.sum()) > int(only_fb.sum())


def _c5_photons(cuda, bunches, nsb, seed):
    """One miniature event of the full-scale stream: ring bunches (ring
    radius U[0.5, 1.3] deg, width 0.1 deg, sizes log-uniform in [1e3, 1e6])
    followed by night-sky photons (uniform within 2 deg, lambda uniform in
    300-600 nm)."""
    from merlict_development_kit_amd import sources
    from merlict_development_kit_amd.bunches import Bunch, generate_bunches
    from merlict_development_kit_amd.propagation import PhotonBatch
    rng = np.random.default_rng(seed)
    radii = rng.uniform(0.5, 1.3, bunches)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(1e6), bunches)).astype(np.int64)
    ring = generate_bunches([Bunch(sources.shower_ring(float(radii[k])), seed=seed * 1000 + k,
                                   num_photons=int(sizes[k])) for k in range(bunches)], device=cuda)
    sky = sources.night_sky().generate(nsb, seed=seed, device=cuda)
    cat = [torch.cat([getattr(ring, f.name), getattr(sky, f.name)]) for f in dataclasses.fields(PhotonBatch)]
    return PhotonBatch(*cat)


@pytest.mark.gpu
def test_gpu_hybrid_shower_and_night_sky_stream_matches_plain(cuda, fact_trained):
    """A miniature of the full-scale hybrid stream: 64 Cherenkov-ring bunches
    and 1e6 night-sky photons through the hybrid (entropy bound off) and
    through the plain propagator, each over 4 propagation seeds; the per-pixel
    and per-pixel-and-time L1 errors against plain stay within stated bars
    above the noise floor (plain against plain under other seeds)."""
    from merlict_development_kit_amd import PropagationConfig, histogram_arrivals
    from merlict_development_kit_amd import surrogate as SG
    gs, desc, sur, dom, cells = fact_trained
    ph = _c5_photons(cuda, 64, 1_000_000, seed=5)
    binning = _binning(dom)
    passes = 4

    def plain(seed0):
        c = binning.empty(cuda)
        for s in range(passes):
            histogram_arrivals(gs.propagate(ph, PropagationConfig(5, seed=seed0 + s)), binning, counts=c)
        return c.cpu().numpy().astype(np.float64)

    ref, again = plain(0), plain(100)
    hy = binning.empty(cuda)
    fallback = 0
    for s in range(passes):
        _, k = SG.propagate_hybrid(gs, sur, dom, cells, ph, binning, PropagationConfig(5, seed=s),
                                   max_entropy=float("inf"), seed=s, counts=hy)
        fallback += k
    hy = hy.cpu().numpy().astype(np.float64)

    def l1(a, b):
        return float(np.abs(a - b).sum() / b.sum())

    pix, pix_t = l1(hy.sum(axis=1), ref.sum(axis=1)), l1(hy, ref)
    floor, floor_t = l1(again.sum(axis=1), ref.sum(axis=1)), l1(again, ref)
    print(f"{len(ph)} photons x {passes} passes, {int(ref.sum())} plain arrivals, fallbacks {fallback}; "
          f"pixel L1 {pix:.4f} (floor {floor:.4f}), pixel-time L1 {pix_t:.4f} (floor {floor_t:.4f}), "
          f"detected hybrid/plain {hy.sum() / ref.sum():.4f}")
    # bars for a surrogate trained on 1e8 photons (measured: pixel L1 0.7 %
    # over a 0.34 % floor, pixel-time L1 2.1 % over 0.66 %)
    assert abs(hy.sum() / ref.sum() - 1) < 0.01
    assert pix < floor + 0.01
    assert pix_t < floor_t + 0.025


@pytest.mark.gpu
@pytest.mark.parametrize("light_guides,mode", [(True, "hexagon"), (True, "nearest"), (False, "hexagon")])
def test_gpu_hybrid_cells_follow_the_fact_pixel_table(cuda, light_guides, mode):
    """The hybrid's cell lookup on the FACT camera, whose pixel table sits off
    a regular lattice by up to 0.4 mm: a network with zero weights whose
    biases put every photon at its nominal image point plus a fixed offset
    (survival logit 20, log sigma -9: a 1.2 um spread) bins each photon into
    the hexagon that holds the point (a light guide's entry, or a bare
    pixel) and nowhere in the gaps between them, or, in the nearest-centre
    mode, into the pixel whose centre is nearest within hexagons that close
    the gaps, found by brute force over all 1440 pixels on the host.  Only photons within 0.6 pitch of a centre
    count, and those within 10 um of a cell boundary are left out of the
    batch, so the histograms agree exactly."""
    from dataclasses import replace

    from scipy.spatial import cKDTree

    import surrogate_oracle as SO
    from merlict_development_kit_amd import GpuScene, PropagationConfig, SceneDescription
    from merlict_development_kit_amd import surrogate as SG
    from merlict_development_kit_amd.propagation import PhotonBatch
    desc = SceneDescription.fact(light_guides=light_guides)
    gs = GpuScene(desc)
    cells = SG.HexCells(desc, cuda, mode)
    assert cells.centres is not None
    r_hex = (5.2772e-3 if light_guides else 4.95e-3) * np.sqrt(3) / 2
    if mode == "hexagon":
        assert cells.inner_radius == pytest.approx(r_hex, rel=1e-4)
    else:
        assert 4.9e-3 < cells.inner_radius < 5.0e-3  # half the widest row spacing of the FACT table
    dom = replace(SG.fact_domain(), image_scale=4.889, position_scale=0.01, time_offset=2e-8, time_scale=1e-9)
    offset = (0.25, -0.1875)  # of the image point, in units of position_scale (exact in bf16)
    p = np.zeros(SO.NUM_PARAMS)
    b_last = SO.NUM_PARAMS - 57
    p[b_last] = 20.0
    for q in range(8):
        p[b_last + 2 + 7 * q:b_last + 8 + 7 * q] = (offset[0], offset[1], 0.5, -9.0, -9.0, -9.0)
    sur = SG.Surrogate(seed=0)
    sur.set_params(torch.tensor(p, dtype=torch.float32, device=cuda))
    ph = dom.source(10.0).generate(400_000, seed=17, device=cuda)
    f, _ = SG.prepare(dom, ph)
    f = f.cpu().numpy().astype(np.float64)
    sa = np.sin(dom.max_angle)
    dx, dy = f[:, 2] * sa, f[:, 3] * sa
    dz = -np.sqrt(np.maximum(1 - dx * dx - dy * dy, 1e-12))
    x = dom.image_scale * dx / dz + offset[0] * dom.position_scale
    y = dom.image_scale * dy / dz + offset[1] * dom.position_scale
    centres = cells.centres.cpu().numpy().reshape(-1, 2).astype(np.float64)
    dist, near = cKDTree(centres).query(np.stack([x, y], axis=1), k=3)
    # the nearest centre whose hexagon holds the point
    ang = np.radians([0.0, 60.0, 120.0])
    normals = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    cell = np.full(len(x), -1)
    margin = np.full(len(x), np.inf)
    for k in range(3):
        rel = np.stack([x - centres[near[:, k], 0], y - centres[near[:, k], 1]], axis=1)
        depth = cells.inner_radius - np.abs(rel @ normals.T).max(axis=1)  # > 0 inside
        cell = np.where((depth > 0) & (cell < 0), near[:, k], cell)
        margin = np.minimum(margin, np.abs(depth))
    if mode == "nearest":
        # and the distance to the boundary between the two nearest centres
        margin = np.minimum(margin, (dist[:, 1] ** 2 - dist[:, 0] ** 2) / (
            2 * np.linalg.norm(centres[near[:, 1]] - centres[near[:, 0]], axis=1)))
    keep = (dist[:, 0] < 0.6 * cells.pitch) & (margin > 1e-5)
    print(f"{keep.sum()} of {len(x)} photons kept, {int((cell[keep] >= 0).sum())} on "
          f"{len(np.unique(cell[keep][cell[keep] >= 0]))} cells")
    assert len(np.unique(cell[keep][cell[keep] >= 0])) > 1300
    if mode == "hexagon":
        assert (cell[keep] < 0).sum() > 100  # photons between the hexagons
    else:
        # no gaps inside the camera; points without a cell lie off its rim
        inner = np.hypot(centres[near[:, 0], 0], centres[near[:, 0], 1]) < 0.17
        assert ((cell < 0) & keep & inner).sum() == 0 and ((cell < 0) & keep).sum() > 100
    idx = torch.tensor(np.nonzero(keep)[0], device=cuda)
    sub = PhotonBatch(*[getattr(ph, n).index_select(0, idx).contiguous() for n in (
        "support_x", "support_y", "support_z", "direction_x", "direction_y", "direction_z", "wavelength")])
    binning = _binning(dom)
    got, k = SG.propagate_hybrid(gs, sur, dom, cells, sub, binning, PropagationConfig(5), max_entropy=float("inf"),
                                 seed=1)
    assert k == 0
    want = np.zeros((1440, 16), np.int64)
    on = cell[keep] >= 0
    np.add.at(want, (cell[keep][on], 8), 1)
    got = got.cpu().numpy().astype(np.int64)
    print(f"binned {got.sum()}, expected {want.sum()}")
    assert np.array_equal(got, want)
