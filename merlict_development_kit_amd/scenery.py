"""Scene descriptions: the flat frame list the propagator consumes.

A :class:`SceneDescription` owns a native ``mlgpu_desc`` (built by the C++
builders or loaded from a ``MLGPUSC1`` file).  It mirrors merlict's
``Scenery`` (merlict/scenery/Scenery.h:12-21): a root frame with children,
the surface functions and the sensor ids.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class SceneDescription:
    def __init__(self, handle: C.c_void_p):
        self._h = handle
        self._view = N.SceneDesc()
        N.check(N.lib().mlgpu_desc_view(self._h, C.byref(self._view)), "mlgpu_desc_view")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().mlgpu_desc_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    # ------------------------------------------------------------ builders
    @classmethod
    def lens(cls, focal_length=1.0, aperture_radius=0.25, refractive_index=1.49, cell_pitch=0.06):
        """Bi-convex lens test bench: merlict BiConvexLens + 19 hexagonal cells."""
        h = C.c_void_p()
        N.check(N.lib().mlgpu_build_lens_scene(focal_length, aperture_radius, refractive_index, cell_pitch,
                                                C.byref(h)), "mlgpu_build_lens_scene")
        return cls(h)

    @classmethod
    def fact(cls, light_guides: bool = True, focal_plane: bool = False):
        """FACT telescope as merlict's FACT propagation app builds it
        (fact_propagation.cpp:282-341): 30-facet Davies-Cotton dish with
        FACT's mirror reflectivity, the 1440 pixels of FACT's pixel table
        (sensor id = chid), with PMMA light guides in front of the pixels or
        bare pixels.
        focal_plane: the same dish with one sensor disc at the light guides'
        entrance plane instead of the camera (mlgpu_build_fact_scene)."""
        h = C.c_void_p()
        camera = 2 if focal_plane else (1 if light_guides else 0)
        N.check(N.lib().mlgpu_build_fact_scene(camera, C.byref(h)), "mlgpu_build_fact_scene")
        return cls(h)

    @classmethod
    def plenoscope(cls, focal_length=106.5, max_outer_aperture_radius=35.5, min_inner_aperture_radius=3.5,
                   facet_inner_hex_radius=0.75, gap_between_facets=0.02, max_FoV_diameter_deg=6.5,
                   hex_pixel_FoV_flat2flat_deg=0.06667, num_paxel_on_pixel_diagonal=9, housing_overhead=1.1,
                   lens_refraction=1.49, mirror_reflection=0.8):
        """The plenoscope of merlict's calibration example scenery
        (merlict_portal_plenoscope/apps/examples/PlenoscopeCalibration/scenery/scenery.json)."""
        h = C.c_void_p()
        N.check(N.lib().mlgpu_build_plenoscope_scene(
            focal_length, max_outer_aperture_radius, min_inner_aperture_radius, facet_inner_hex_radius,
            gap_between_facets, max_FoV_diameter_deg, hex_pixel_FoV_flat2flat_deg, num_paxel_on_pixel_diagonal,
            housing_overhead, lens_refraction, mirror_reflection, C.byref(h)), "mlgpu_build_plenoscope_scene")
        return cls(h)

    @classmethod
    def from_arrays(cls, mother, pos, rot_mode, rot, kind, params, surface=None, allowed_frame=None,
                    sensor_id=None, functions=()):
        """A description from per-frame arrays (frame 0 the root, mothers
        before children): pos (n, 3), rot (n, 4) as xyz angles (rot_mode 0)
        or axis + angle (1), params (n, 6), surface (n, 6) function indices
        (-1: merlict's defaults), and functions as (x, y) tables."""
        n = len(mother)
        mother = np.ascontiguousarray(mother, np.int32)
        pos = np.ascontiguousarray(np.reshape(pos, (n, 3)), np.float64)
        rot_mode = np.ascontiguousarray(rot_mode, np.int32)
        rot = np.ascontiguousarray(np.reshape(rot, (n, 4)), np.float64)
        kind = np.ascontiguousarray(kind, np.int32)
        params = np.ascontiguousarray(np.reshape(params, (n, N.MLGPU_NUM_PARAMS)), np.float64)
        surface = np.full((n, N.MLGPU_NUM_SURFACE_SLOTS), -1, np.int32) if surface is None else \
            np.ascontiguousarray(np.reshape(surface, (n, N.MLGPU_NUM_SURFACE_SLOTS)), np.int32)
        allowed = np.full(n, -1, np.int32) if allowed_frame is None else np.ascontiguousarray(allowed_frame, np.int32)
        sensor = np.full(n, -1, np.int32) if sensor_id is None else np.ascontiguousarray(sensor_id, np.int32)
        starts = np.cumsum([0] + [len(x) for x, _ in functions]).astype(np.int32)
        fx = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x, _ in functions])
                                  if functions else np.zeros(1), np.float64)
        fy = np.ascontiguousarray(np.concatenate([np.asarray(y, np.float64) for _, y in functions])
                                  if functions else np.zeros(1), np.float64)
        p = lambda a, t: a.ctypes.data_as(C.POINTER(t))  # noqa: E731
        v = N.SceneDesc(n, p(mother, C.c_int32), p(pos, C.c_double), p(rot_mode, C.c_int32), p(rot, C.c_double),
                        p(kind, C.c_int32), p(params, C.c_double), p(surface, C.c_int32), p(allowed, C.c_int32),
                        p(sensor, C.c_int32), len(functions), p(starts, C.c_int32), p(fx, C.c_double),
                        p(fy, C.c_double))
        h = C.c_void_p()
        N.check(N.lib().mlgpu_desc_create(C.byref(v), C.byref(h)), "mlgpu_desc_create")
        return cls(h)

    @classmethod
    def load(cls, path: str):
        h = C.c_void_p()
        N.check(N.lib().mlgpu_desc_load(str(path).encode(), C.byref(h)), "mlgpu_desc_load")
        return cls(h)

    def save(self, path: str) -> None:
        N.check(N.lib().mlgpu_desc_save(self._h, str(path).encode()), "mlgpu_desc_save")

    # ------------------------------------------------------------- access
    @property
    def view(self) -> N.SceneDesc:
        """The C view (valid while this object lives)."""
        return self._view

    @property
    def num_frames(self) -> int:
        return int(self._view.n_frames)

    def arrays(self) -> dict:
        v, n = self._view, self.num_frames
        nf = int(v.n_functions)

        def arr(ptr, count, dtype):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).copy().astype(dtype)

        fs = arr(v.func_start, nf + 1, np.int32)
        npts = int(fs[-1]) if nf > 0 else 0
        return {
            "mother": arr(v.mother, n, np.int32),
            "pos": arr(v.pos, 3 * n, np.float64).reshape(n, 3),
            "rot_mode": arr(v.rot_mode, n, np.int32),
            "rot": arr(v.rot, 4 * n, np.float64).reshape(n, 4),
            "kind": arr(v.kind, n, np.int32),
            "params": arr(v.params, N.MLGPU_NUM_PARAMS * n, np.float64).reshape(n, N.MLGPU_NUM_PARAMS),
            "surface": arr(v.surface, N.MLGPU_NUM_SURFACE_SLOTS * n, np.int32).reshape(n, N.MLGPU_NUM_SURFACE_SLOTS),
            "allowed_frame": arr(v.allowed_frame, n, np.int32),
            "sensor_id": arr(v.sensor_id, n, np.int32),
            "func_start": fs,
            "func_x": arr(v.func_x, npts, np.float64),
            "func_y": arr(v.func_y, npts, np.float64),
        }

    def frame_poses(self) -> np.ndarray:
        """frame2world of every frame (n, 3, 4), as merlict's initialised tree
        has it."""
        out = np.zeros((self.num_frames, 3, 4), dtype=np.float64)
        N.check(N.lib().mlgpu_desc_frame_poses(C.byref(self._view), out.ctypes.data_as(C.POINTER(C.c_double))),
                "mlgpu_desc_frame_poses")
        return out

    def sensor_frames(self) -> np.ndarray:
        """Frame index of every sensor, indexed by sensor id."""
        s = self.arrays()["sensor_id"]
        idx = np.nonzero(s >= 0)[0]
        out = np.full(int(s.max()) + 1 if idx.size else 0, -1, dtype=np.int64)
        out[s[idx]] = idx
        return out

