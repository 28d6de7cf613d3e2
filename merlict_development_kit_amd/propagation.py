"""Device-side propagation: the drop-in for merlict's propagation API.

``propagate_photons_in_frame_with_config`` mirrors
``merlict::propagate_photons_in_frame_with_config`` (merlict/Photons.cpp:9-22):
photons in, per-photon final state out.  Photons live in HBM as a structure of
arrays of float64 (torch tensors on the GPU); the propagation runs in the HIP
kernels of ``libmlgpu.so``.  torch provides device memory and the stream only.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields
from typing import Optional

import torch

from . import _native as N
from .scenery import SceneDescription


@dataclass
class PhotonBatch:
    """Photons as SoA float64 tensors (same device)."""
    support_x: torch.Tensor
    support_y: torch.Tensor
    support_z: torch.Tensor
    direction_x: torch.Tensor
    direction_y: torch.Tensor
    direction_z: torch.Tensor
    wavelength: torch.Tensor

    def __len__(self) -> int:
        return int(self.support_x.numel())

    @property
    def device(self) -> torch.device:
        return self.support_x.device

    def to(self, device) -> "PhotonBatch":
        return PhotonBatch(**{f.name: getattr(self, f.name).to(device) for f in fields(self)})

    def slice(self, a: int, b: int) -> "PhotonBatch":
        return PhotonBatch(**{f.name: getattr(self, f.name)[a:b] for f in fields(self)})

    def numpy(self) -> dict:
        return {f.name: getattr(self, f.name).detach().cpu().numpy() for f in fields(self)}

    @classmethod
    def from_numpy(cls, d: dict, device="cpu") -> "PhotonBatch":
        return cls(**{f.name: torch.as_tensor(d[f.name], dtype=torch.float64, device=device).contiguous()
                      for f in fields(cls)})

    def c_struct(self) -> N.Photons:
        p = N.Photons()
        for f in fields(self):
            t = getattr(self, f.name)
            if t.dtype != torch.float64 or not t.is_contiguous():
                raise ValueError(f"{f.name} must be a contiguous float64 tensor")
            setattr(p, f.name, t.data_ptr())
        return p


@dataclass
class PropagationResult:
    """Per-photon final state (merlict Photon accessors, RayForPropagation.h:53-59)."""
    final_frame: torch.Tensor        # int32, frame index, -1 = void
    sensor_id: torch.Tensor          # int32, -1 = not on a sensor
    final_interaction: torch.Tensor  # uint8, merlict::Interaction
    num_interactions: torch.Tensor   # uint8, history length incl. production
    error: torch.Tensor              # uint8
    time_of_flight: torch.Tensor     # float64, seconds
    x: torch.Tensor                  # float32, object frame
    y: torch.Tensor
    cx: torch.Tensor                 # float32, -incident direction, object frame
    cy: torch.Tensor
    accumulated_distance: torch.Tensor  # float64, metres (RayForPropagation.cpp:91-98)

    @classmethod
    def empty(cls, n: int, device) -> "PropagationResult":
        i32 = dict(dtype=torch.int32, device=device)
        u8 = dict(dtype=torch.uint8, device=device)
        f32 = dict(dtype=torch.float32, device=device)
        return cls(final_frame=torch.empty(n, **i32), sensor_id=torch.empty(n, **i32),
                   final_interaction=torch.empty(n, **u8), num_interactions=torch.empty(n, **u8),
                   error=torch.empty(n, **u8), time_of_flight=torch.empty(n, dtype=torch.float64, device=device),
                   x=torch.empty(n, **f32), y=torch.empty(n, **f32), cx=torch.empty(n, **f32),
                   cy=torch.empty(n, **f32), accumulated_distance=torch.empty(n, dtype=torch.float64, device=device))

    def c_struct(self) -> N.Results:
        r = N.Results()
        for f in fields(self):
            setattr(r, f.name, getattr(self, f.name).data_ptr())
        return r

    def numpy(self) -> dict:
        return {f.name: getattr(self, f.name).detach().cpu().numpy() for f in fields(self)}


@dataclass
class PropagationConfig:
    """merlict::PropagationConfig (merlict/PropagationConfig.cpp:7) plus the Philox key."""
    max_num_interactions_per_photon: int = 5
    seed: int = 0
    photon_id_offset: int = 0

    def c_struct(self) -> N.Config:
        return N.Config(self.max_num_interactions_per_photon, self.seed, self.photon_id_offset)


def _stream_ptr(stream: Optional[torch.cuda.Stream]) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class GpuScene:
    """A scene flattened and resident on the current GPU (mlgpu_scene_create)."""

    def __init__(self, desc: SceneDescription):
        if not torch.cuda.is_available():
            raise RuntimeError("GpuScene needs a GPU")
        self.desc = desc
        self._h = C.c_void_p()
        N.check(N.lib().mlgpu_scene_create(C.byref(desc.view), C.byref(self._h)), "mlgpu_scene_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().mlgpu_scene_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        """The mlgpu_scene handle."""
        return self._h

    _SCHEDULES = {"auto": N.SCHEDULE_AUTO, "fused": N.SCHEDULE_FUSED, "wavefront": N.SCHEDULE_WAVEFRONT,
                  "persistent": N.SCHEDULE_PERSISTENT}

    def tuning(self) -> dict:
        """Launch parameters (mlgpu_tuning): schedule, trace_budget, blocks_per_cu, tail_level, shade_lanes."""
        t = N.Tuning()
        N.check(N.lib().mlgpu_scene_get_tuning(self._h, C.byref(t)), "mlgpu_scene_get_tuning")
        names = {v: k for k, v in self._SCHEDULES.items()}
        return {"schedule": names[t.schedule], "trace_budget": t.trace_budget, "blocks_per_cu": t.blocks_per_cu,
                "tail_level": t.tail_level, "shade_lanes": t.shade_lanes}

    def set_tuning(self, schedule: Optional[str] = None, trace_budget: Optional[int] = None,
                   blocks_per_cu: Optional[int] = None, tail_level: Optional[int] = None,
                   shade_lanes: Optional[int] = None) -> None:
        """Change launch parameters; results do not depend on them.  schedule is
        'auto', 'fused' (one lane per photon), 'wavefront' (trace/shade per
        bounce level) or 'persistent' (resident waves, a lane takes the next
        photon when its photon ends); tail_level (wavefront) is the level from
        which a photon's remaining bounces run in one lane (<= 0: never);
        shade_lanes (persistent) is how many lanes of a wave wait with ended
        traces before the wave shades them."""
        cur = self.tuning()
        t = N.Tuning(self._SCHEDULES[schedule if schedule is not None else cur["schedule"]],
                     trace_budget if trace_budget is not None else cur["trace_budget"],
                     blocks_per_cu if blocks_per_cu is not None else cur["blocks_per_cu"],
                     tail_level if tail_level is not None else cur["tail_level"],
                     shade_lanes if shade_lanes is not None else cur["shade_lanes"])
        N.check(N.lib().mlgpu_scene_set_tuning(self._h, C.byref(t)), "mlgpu_scene_set_tuning")

    def set_schedule(self, schedule: str) -> None:
        self.set_tuning(schedule=schedule)

    def autotune(self, photons: PhotonBatch, config: PropagationConfig = PropagationConfig(), repeats: int = 3,
                 world_frame: int = 0, stream: Optional[torch.cuda.Stream] = None):
        """Time the candidate tunings on these photons (mlgpu_scene_autotune),
        keep the fastest; returns (tuning, mean ms per call)."""
        if photons.device.type != "cuda":
            raise ValueError("photons must be on the GPU")
        ph, cfg = photons.c_struct(), config.c_struct()
        t, ms = N.Tuning(), C.c_double()
        N.check(N.lib().mlgpu_scene_autotune(self._h, world_frame, C.byref(ph), len(photons), C.byref(cfg), repeats,
                                             C.byref(t), C.byref(ms), _stream_ptr(stream)), "mlgpu_scene_autotune")
        return self.tuning(), ms.value

    def stats(self) -> dict:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        N.check(N.lib().mlgpu_scene_stats(self._h, C.byref(a), C.byref(b), C.byref(c)), "mlgpu_scene_stats")
        return {"nodes": a.value, "leaves": b.value, "device_bytes": c.value}

    def propagate(self, photons: PhotonBatch, config: PropagationConfig = PropagationConfig(),
                  world_frame: int = 0, out: Optional[PropagationResult] = None,
                  stream: Optional[torch.cuda.Stream] = None) -> PropagationResult:
        n = len(photons)
        if photons.device.type != "cuda":
            raise ValueError("photons must be on the GPU")
        if out is None:
            out = PropagationResult.empty(n, photons.device)
        ph, rs, cfg = photons.c_struct(), out.c_struct(), config.c_struct()
        N.check(N.lib().mlgpu_propagate(self._h, world_frame, C.byref(ph), n, C.byref(cfg), C.byref(rs),
                                        _stream_ptr(stream)), "mlgpu_propagate")
        return out

    def propagate_history(self, photons: PhotonBatch, config: PropagationConfig = PropagationConfig(),
                          world_frame: int = 0, stream: Optional[torch.cuda.Stream] = None):
        """propagate() that also records each photon's history after its
        production (mlgpu_propagate_history): returns (result, history) with
        history = {"bounces": uint8 (n, capacity, 64) mlgpu_bounce records
        (decode_bounces reads them), "num_bounces": int32 (n),
        "final_ray": float64 (n, 6) last support and direction}."""
        n = len(photons)
        if photons.device.type != "cuda":
            raise ValueError("photons must be on the GPU")
        dev = photons.device
        cap = max(1, int(config.max_num_interactions_per_photon))
        out = PropagationResult.empty(n, dev)
        hist = {
            "bounces": torch.empty((n, cap, 64), dtype=torch.uint8, device=dev),
            "num_bounces": torch.empty(n, dtype=torch.int32, device=dev),
            "final_ray": torch.empty((n, 6), dtype=torch.float64, device=dev),
        }
        h = N.History(hist["bounces"].data_ptr(), cap, hist["num_bounces"].data_ptr(), hist["final_ray"].data_ptr())
        ph, rs, cfg = photons.c_struct(), out.c_struct(), config.c_struct()
        N.check(N.lib().mlgpu_propagate_history(self._h, world_frame, C.byref(ph), n, C.byref(cfg), C.byref(rs),
                                                C.byref(h), _stream_ptr(stream)), "mlgpu_propagate_history")
        return out, hist

    def count_errors(self, result: PropagationResult, stream=None) -> int:
        rs = result.c_struct()
        k = C.c_int64()
        N.check(N.lib().mlgpu_count_errors(C.byref(rs), len(result.error), _stream_ptr(stream), C.byref(k)),
                "mlgpu_count_errors")
        return k.value

    def first_hit(self, rays: PhotonBatch, root_frame: int = 0, stream=None) -> dict:
        n = len(rays)
        dev = rays.device
        out = {
            "frame": torch.empty(n, dtype=torch.int32, device=dev),
            "distance": torch.empty(n, dtype=torch.float64, device=dev),
            "position_x": torch.empty(n, dtype=torch.float64, device=dev),
            "position_y": torch.empty(n, dtype=torch.float64, device=dev),
            "position_z": torch.empty(n, dtype=torch.float64, device=dev),
        }
        h = N.Hits(*(out[k].data_ptr() for k in ("frame", "distance", "position_x", "position_y", "position_z")))
        r = rays.c_struct()
        N.check(N.lib().mlgpu_first_hit(self._h, root_frame, C.byref(r), n, C.byref(h), _stream_ptr(stream)),
                "mlgpu_first_hit")
        return out


def propagate_photons_in_frame_with_config(photons: PhotonBatch, world: GpuScene,
                                           settings: PropagationConfig = PropagationConfig(),
                                           world_frame: int = 0) -> PropagationResult:
    """Drop-in of merlict::propagate_photons_in_frame_with_config (merlict/Photons.cpp:9-22).

    Raises like the reference when a photon has lambda <= 0
    (std::invalid_argument, Photon.cpp:20-28 -> ValueError) or a wavelength
    outside a surface table (std::out_of_range, Limits.cpp:37-46 -> IndexError).
    """
    res = world.propagate(photons, settings, world_frame=world_frame)
    if world.count_errors(res):
        err = res.error.cpu()
        if bool((err == N.PHOTON_BAD_WAVELENGTH).any()):
            raise ValueError("photon wavelength must be > 0")
        raise IndexError("photon wavelength outside a surface function's range")
    return res


def decode_bounces(bounces: torch.Tensor):
    """mlgpu_bounce records (uint8, last dimension 64) as a numpy structured
    array with fields position (3), normal (3), distance, frame, interaction,
    from_outside."""
    import numpy as np
    raw = bounces.detach().cpu().contiguous().numpy()
    if raw.shape[-1] != 64:
        raise ValueError("expected 64-byte mlgpu_bounce records")
    return raw.view(np.dtype(N.BOUNCE_FIELDS))[..., 0]
