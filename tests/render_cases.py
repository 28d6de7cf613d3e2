"""Scenes, cameras and colours of the rendering tests and golden renders."""
import math

import numpy as np

import scene_zoo
from merlict_development_kit_amd import SceneDescription
from merlict_development_kit_amd import visual as V


def render_cases():
    """(name, description, camera, global illumination, seed): the primitive
    zoo, the FACT telescope from above, the lens bench without global
    illumination."""
    return [
        ("zoo", scene_zoo.zoo(), V.PinHoleCamera((0.0, -3.0, 3.0), (math.radians(135), 0.0, 0.0), math.radians(60),
                                                 80, 60), True, 5),
        ("fact", SceneDescription.fact(light_guides=True),
         V.PinHoleCamera((0.0, -6.0, 10.0), (math.radians(149), 0.0, math.radians(10)), math.radians(40), 80, 60),
         True, 6),
        ("lens", SceneDescription.lens(), V.PinHoleCamera((0.0, -1.5, 1.5), (math.radians(135), 0.0, 0.0),
                                                          math.radians(30), 64, 48), False, 7),
    ]


def render_colors(n, seed=1):
    """Outer and inner RGB per frame, uniform in [0, 255)."""
    return np.random.default_rng(seed).uniform(0, 255, (n, 6)).astype(np.float32)


def one_disc(reflectivity: float = 0.0, radius: float = 0.5) -> SceneDescription:
    """One disc (radius r) in the xy plane at the origin, opaque, reflecting
    `reflectivity` on its outer (+z) side, nothing on its inner side."""
    funcs = [((200e-9, 1200e-9), (reflectivity, reflectivity)), ((200e-9, 1200e-9), (0.0, 0.0))]
    return SceneDescription.from_arrays(
        mother=[-1, 0], pos=[(0, 0, 0), (0, 0, 0)], rot_mode=[0, 0], rot=[(0, 0, 0, 0), (0, 0, 0, 0)],
        kind=[0, 5], params=[[0] * 6, [radius, 0, 0, 0, 0, 0]],
        surface=[[-1] * 6, [0, 1, -1, -1, -1, -1]], functions=funcs)
