"""Image rendering on the GPU (mlgpu_render_pinhole) against the reference's
merlict_visual PinHoleCamera + Tracer (golden renders under the same Philox
streams) and the oracle."""
import math

import numpy as np
import pytest
import torch

import pyoracle as O
from merlict_development_kit_amd import GpuScene, SceneDescription
from merlict_development_kit_amd import visual as V
from merlict_development_kit_amd._native import MlgpuError
from render_cases import one_disc, render_cases, render_colors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def test_gpu_render_matches_reference(cuda, golden):
    """The zoo, the FACT telescope and the lens bench: every pixel's
    colour equals the reference Tracer's."""
    g = golden("render")
    for name, desc, cam, gi, seed in render_cases():
        img = V.render(GpuScene(desc), cam, V.RenderConfig(global_illumination=gi), render_colors(desc.num_frames),
                       seed=seed)
        np.testing.assert_array_equal(img.cpu().numpy(), g[name], err_msg=name)


def test_gpu_render_matches_oracle_large(cuda):
    """640 x 480 renders of the zoo and the FACT telescope from other
    cameras: GPU and oracle agree on every pixel."""
    zoo_desc = render_cases()[0][1]
    fact_desc = SceneDescription.fact(light_guides=True)
    cases = [(zoo_desc, V.PinHoleCamera((4.0, -2.0, 2.5), (math.radians(130), 0.0, math.radians(40)),
                                        math.radians(70), 640, 480), True, 11),
             (fact_desc, V.PinHoleCamera((1.0, 1.0, 4.0), (math.radians(180), 0.0, 0.0), math.radians(50), 640, 480),
              True, 12),
             (fact_desc, V.PinHoleCamera((0.3, -0.2, -3.0), (0.0, 0.0, 0.0), math.radians(60), 320, 240), False, 13)]
    for desc, cam, gi, seed in cases:
        cfg = V.RenderConfig(global_illumination=gi)
        col = render_colors(desc.num_frames, seed=seed)
        img = V.render(GpuScene(desc), cam, cfg, col, seed=seed).cpu().numpy()
        ref, err = O.OracleScene(desc).render_pinhole(cam.to_c(), cfg.to_c(), col, seed)
        assert err == 0
        np.testing.assert_array_equal(img, ref)


def test_gpu_render_errors_and_defaults(cuda):
    gs = GpuScene(one_disc(0.0))
    cam = V.PinHoleCamera((0.0, 0.0, 4.0), (math.pi, 0.0, 0.0), math.radians(30), 16, 12)
    img = V.render(gs, cam, V.RenderConfig(False))  # default colours: dark gray
    assert (img[6, 8].cpu().numpy() == np.float32(25.0)).all()
    with pytest.raises(MlgpuError):
        V.render(gs, V.PinHoleCamera((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 0.0, 16, 12))
    with pytest.raises(ValueError):
        V.render(gs, cam, frame_colors=np.zeros((5, 6), np.float32))
    # a reflectivity table that does not cover 533 nm: the reference throws
    short = SceneDescription.from_arrays(
        mother=[-1, 0], pos=[(0, 0, 0), (0, 0, 0)], rot_mode=[0, 0], rot=[(0, 0, 0, 0)] * 2, kind=[0, 5],
        params=[[0] * 6, [0.5, 0, 0, 0, 0, 0]], surface=[[-1] * 6, [0, -1, -1, -1, -1, -1]],
        functions=[((200e-9, 400e-9), (0.5, 0.5))])
    with pytest.raises(ValueError):
        V.render(GpuScene(short), cam)
