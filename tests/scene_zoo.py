"""A scene holding every primitive kind, for first-hit and propagation
checks of the kinds the telescope scenes do not use: caps with hexagonal,
circular and rectangular bounds, plane, disc, hexagonal plane, annulus,
dual-spherical plane, cylinder, triangle, sphere and elliptical caps with
hexagonal bounds (one whose outer radius is restricted to its z curvature
radius), in both rotation modes,
a nested frame, a group of more than 16 children (merlict's octant
clustering) and coplanar overlapping surfaces (exact distance ties)."""
from __future__ import annotations

import numpy as np

from merlict_development_kit_amd import SceneDescription

FRAME, CAP_HEX, CAP_CYL, CAP_RECT, PLANE, DISC, HEX, ANNULUS, DUAL, CYL, TRI, SPHERE, ELLIP_CAP = range(13)


def zoo(functions=None) -> SceneDescription:
    """The primitive zoo; `functions` replaces its two tables (function 0:
    the mirrors' reflectivity, 1: the glass' refractive index)."""
    rows = []  # mother, pos, rot_mode, rot, kind, params, surface

    def add(mother, pos, kind, params=(), rot=(0, 0, 0, 0), mode=0, surface=None):
        prm = list(params) + [0.0] * (6 - len(params))
        rows.append((mother, pos, mode, rot, kind, prm, surface if surface is not None else [-1] * 6))
        return len(rows) - 1

    mirror = [0, 0, -1, -1, -1, -1]      # outer reflection 0.9 (function 0)
    glass = [-1, -1, 1, 1, -1, -1]       # refraction 1.5 both sides (function 1)
    root = add(-1, (0, 0, 0), FRAME)
    add(root, (0.0, 0.0, 0.0), CAP_HEX, (2.0, 0.3), rot=(0.1, 0.2, 0.0), surface=mirror)
    add(root, (1.0, 0.0, 0.0), CAP_CYL, (1.5, 0.25), rot=(0.0, 0.0, 1.0, 0.3), mode=1, surface=mirror)
    add(root, (2.0, 0.0, 0.0), CAP_RECT, (3.0, 0.4, 0.3), rot=(-0.15, 0.05, 0.4))
    add(root, (0.0, 1.0, 0.0), PLANE, (0.5, 0.2), rot=(1.0, 1.0, 0.0, 0.5), mode=1)
    add(root, (1.0, 1.0, 0.1), DISC, (0.2,))
    add(root, (2.0, 1.0, 0.0), HEX, (0.25,), rot=(0.0, 0.3, 0.2))
    add(root, (0.0, 2.0, 0.0), ANNULUS, (0.3, 0.1))
    add(root, (1.0, 2.0, 0.0), DUAL, (0.05, 0.3), rot=(0.2, 0.0, 0.0), surface=glass)
    add(root, (2.0, 2.0, 0.0), CYL, (0.1, 0.5), rot=(1.2, 0.0, 0.0), surface=mirror)
    add(root, (0.0, 3.0, 0.0), TRI, (-0.2, -0.1, 0.25, -0.15, 0.0, 0.3), rot=(0.1, -0.1, 0.7))
    add(root, (1.0, 3.0, 0.0), SPHERE, (0.2,), surface=glass)
    # elliptical caps: focal length, x/y curvature ratio, hex rotation, outer
    # hex radius; the second one's 0.5 m exceeds its z curvature radius
    add(root, (1.0, 4.0, 0.0), ELLIP_CAP, (4.0, 1.2, 0.3, 0.3), rot=(0.1, -0.2, 0.3), surface=glass)
    add(root, (2.0, 4.0, 0.0), ELLIP_CAP, (0.8, 1.3, 0.0, 0.5), rot=(0.0, 1.0, 0.0, 0.4), mode=1, surface=mirror)
    nest = add(root, (2.0, 3.0, 0.2), FRAME, rot=(0.0, 0.0, 0.5))
    add(nest, (0.05, 0.0, 0.0), DISC, (0.15,))
    add(nest, (-0.05, 0.0, -0.1), SPHERE, (0.08,), surface=glass)
    # more than 16 children: merlict clusters them into octant frames
    group = add(root, (3.2, 1.5, 0.0), FRAME)
    rng = np.random.default_rng(5)
    for k in range(24):
        add(group, tuple(rng.uniform(-0.6, 0.6, 3) * (1, 1, 0.2)), HEX if k % 2 else DISC, (0.09,),
            rot=tuple(rng.uniform(-0.3, 0.3, 3)) + (0.0,))
    # coplanar overlap: an annulus and discs in the same plane
    add(root, (-1.0, 1.5, -0.2), ANNULUS, (0.5, 0.2))
    for x, y in ((-1.35, 1.5), (-0.65, 1.5), (-1.0, 1.85)):
        add(root, (x, y, -0.2), DISC, (0.12,))
    n = len(rows)
    return SceneDescription.from_arrays(
        mother=[r[0] for r in rows], pos=[r[1] for r in rows], rot_mode=[r[2] for r in rows],
        rot=[list(r[3]) + [0.0] * (4 - len(r[3])) for r in rows], kind=[r[4] for r in rows],
        params=[r[5] for r in rows], surface=[r[6] for r in rows], sensor_id=list(range(n)),
        functions=functions or [((200e-9, 1200e-9), (0.9, 0.9)), ((200e-9, 1200e-9), (1.5, 1.5))])


def probe_rays(n: int, seed: int) -> dict:
    """Rays from all around towards random points of the zoo."""
    rng = np.random.default_rng(seed)
    centres = np.array([(x, y, 0.0) for y in range(4) for x in range(3)] + [(3.2, 1.5, 0.0), (-1.0, 1.5, -0.2)] +
                       [(1.0, 4.0, 0.0), (2.0, 4.0, 0.2)])
    target = centres[rng.integers(0, len(centres), n)] + rng.uniform(-1, 1, (n, 3)) * (0.4, 0.4, 0.15)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    start = target + 5.0 * u
    d = target - start
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return {"support_x": start[:, 0], "support_y": start[:, 1], "support_z": start[:, 2], "direction_x": d[:, 0],
            "direction_y": d[:, 1], "direction_z": d[:, 2], "wavelength": np.full(n, 433e-9)}
