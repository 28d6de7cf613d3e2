"""mlgpu_propagate_history: the per-bounce history it records agrees with the
final state mlgpu_propagate reports, photon by photon (same Philox streams).

What merlict's RayForPropagation derives from its history
(RayForPropagation.cpp:91-136, Photon.cpp:44-56) is recomputed from the
recorded bounces and compared with the propagator's own outputs:
  * the results equal mlgpu_propagate's bit for bit;
  * num_bounces = num_interactions - 1; the last bounce carries the final
    frame and interaction;
  * accumulated_distance = the sum of the bounce distances (same order);
  * the last bounce's object-frame x, y are the reported x, y;
  * a photon absorbed in void ends 1e4 m along its last ray.
The drop-in comparison with merlict itself is tests/test_facade.py.
"""
import numpy as np
import pytest
import torch

import photon_sets as sources
import scene_zoo
from merlict_development_kit_amd import GpuScene, PhotonBatch, PropagationConfig, SceneDescription
from merlict_development_kit_amd.propagation import decode_bounces

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.mark.parametrize("scene", ["zoo", "fact_lg", "lens"])
def test_history_agrees_with_final_state(cuda, scene):
    print("history", scene, flush=True)
    if scene == "zoo":
        desc = scene_zoo.zoo()
        ph = PhotonBatch.from_numpy(scene_zoo.probe_rays(40_000, 5))
        cfg = PropagationConfig(max_num_interactions_per_photon=12, seed=9)
    elif scene == "fact_lg":
        desc = SceneDescription.fact(light_guides=True)
        ph = sources.fact_bunch(40_000, seed=6)
        cfg = PropagationConfig(max_num_interactions_per_photon=8, seed=9)
    else:
        desc = SceneDescription.lens()
        ph = sources.lens_beam(40_000, seed=7)
        cfg = PropagationConfig(max_num_interactions_per_photon=5, seed=9)
    ph = ph.to(cuda)
    gs = GpuScene(desc)
    plain = gs.propagate(ph, cfg).numpy()
    res, hist = gs.propagate_history(ph, cfg)
    torch.cuda.synchronize()
    r = res.numpy()
    for k in plain:
        np.testing.assert_array_equal(r[k], plain[k], err_msg=k)

    ok = r["error"] == 0
    assert ok.mean() > 0.99
    nb = hist["num_bounces"].cpu().numpy()
    np.testing.assert_array_equal(nb[ok], r["num_interactions"][ok].astype(np.int32) - 1)
    b = decode_bounces(hist["bounces"])
    ray = hist["final_ray"].cpu().numpy()

    idx = np.nonzero(ok & (nb > 0))[0]
    last = b[idx, nb[idx] - 1]
    np.testing.assert_array_equal(last["frame"], r["final_frame"][idx])
    np.testing.assert_array_equal(last["interaction"], r["final_interaction"][idx])
    np.testing.assert_array_equal(last["position"][:, 0].astype(np.float32), r["x"][idx])
    np.testing.assert_array_equal(last["position"][:, 1].astype(np.float32), r["y"][idx])

    # accumulated distance: the running sum, in bounce order
    cap = b.shape[1]
    dist = np.where(np.arange(cap)[None, :] < nb[:, None], b["distance"], 0.0)
    acc = np.zeros(len(nb))
    for k in range(cap):
        acc = acc + dist[:, k]
    np.testing.assert_array_equal(acc[ok], r["accumulated_distance"][ok])

    # void: the last point is 1e4 m along the last ray (RayAndFrame.cpp:160-167)
    void = idx[last["frame"] == -1]
    if len(void):
        lv = b[void, nb[void] - 1]
        np.testing.assert_array_equal(lv["distance"], 1e4)
        expect = ray[void, :3] + ray[void, 3:] * 1e4
        np.testing.assert_array_equal(lv["position"], expect)
        np.testing.assert_array_equal(lv["normal"], ray[void, 3:])
    # every recorded frame is a real frame or void, every side flag 0/1
    rec = np.arange(cap)[None, :] < nb[:, None]
    assert np.all(b["frame"][rec] >= -1) and np.all(b["frame"][rec] < desc.num_frames)
    assert set(np.unique(b["from_outside"][rec])) <= {0, 1}
    print(scene, "photons", len(nb), "bounces", int(nb[ok].sum()), "void", len(void), flush=True)


def test_history_capacity_checked(cuda):
    gs = GpuScene(SceneDescription.lens())
    ph = sources.lens_beam(16, seed=1).to(cuda)
    import ctypes as C
    from merlict_development_kit_amd import _native as N
    cfg = PropagationConfig(max_num_interactions_per_photon=5)
    res = gs.propagate(ph, cfg)
    bounces = torch.empty((16, 4, 64), dtype=torch.uint8, device=cuda)
    nb = torch.empty(16, dtype=torch.int32, device=cuda)
    h = N.History(bounces.data_ptr(), 4, nb.data_ptr(), None)
    p, rs, c = ph.c_struct(), res.c_struct(), cfg.c_struct()
    rc = N.lib().mlgpu_propagate_history(gs.handle, 0, C.byref(p), 16, C.byref(c), C.byref(rs), C.byref(h), None)
    assert rc == 1  # MLGPU_ERR_ARGUMENT
