"""Image rendering on the GPU (merlict_visual): the pin-hole camera and the
Tracer over a scene's first hits, and the PortablePixMap (P6) format
(include/mlgpu.h)."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from pathlib import Path
from typing import Optional, Sequence

import numpy as np
import torch

from . import _native as N
from .propagation import GpuScene

# merlict/Color.h:19-27
COLOR_WHITE = (255.0, 255.0, 255.0)
COLOR_SKY_BLUE = (127.0, 188.0, 255.0)
COLOR_GRAY = (128.0, 128.0, 128.0)
COLOR_DARK_GRAY = (25.0, 25.0, 25.0)
COLOR_BLACK = (0.0, 0.0, 0.0)
COLOR_RED = (255.0, 0.0, 0.0)
COLOR_GREEN = (0.0, 255.0, 0.0)
COLOR_BLUE = (0.0, 0.0, 255.0)
COLOR_GRASS_GREEN = (22.0, 91.0, 49.0)


@dataclass
class PinHoleCamera:
    """visual::PinHoleCamera after set_pos_rot_fov (PinHoleCamera.cpp:10-50)."""
    position: Sequence[float] = (0.0, 0.0, 0.0)
    rotation: Sequence[float] = (0.0, 0.0, 0.0)   # Tait-Bryan x, y, z angles (Rot3)
    field_of_view: float = math.radians(60.0)
    num_cols: int = 128
    num_rows: int = 72

    def to_c(self) -> N.PinholeCamera:
        return N.PinholeCamera((C.c_double * 3)(*self.position), (C.c_double * 3)(*self.rotation),
                               self.field_of_view, self.num_cols, self.num_rows)


@dataclass
class RenderConfig:
    """The parts of visual::Config the Tracer reads (Config.cpp:6-24), plus
    the bound on reflections and refractions per ray (the Config's
    max_interaction_depth, 42)."""
    global_illumination: bool = True
    incoming_direction: Sequence[float] = (0.6, 0.2, 1.0)
    sky_color: Sequence[float] = COLOR_SKY_BLUE
    max_interaction_depth: int = 42

    def to_c(self) -> N.RenderConfig:
        return N.RenderConfig(int(bool(self.global_illumination)), self.max_interaction_depth,
                              (C.c_double * 3)(*self.incoming_direction), (C.c_float * 3)(*self.sky_color), 0.0)


def render(scene: GpuScene, camera: PinHoleCamera, config: RenderConfig = RenderConfig(),
           frame_colors: Optional[np.ndarray] = None, seed: int = 0, world_frame: int = 0,
           stream=None) -> torch.Tensor:
    """The image (num_rows, num_cols, 3) float32 on the GPU, colours 0..255
    (mlgpu_render_pinhole); frame_colors: (num_frames, 6) outer then inner
    RGB per frame of the scene's description (None: merlict's dark gray).
    Raises ValueError when a ray met a surface function that does not cover
    533 nm (the reference throws std::out_of_range)."""
    n_frames = scene.desc.num_frames
    cols = None
    if frame_colors is not None:
        cols = np.ascontiguousarray(frame_colors, np.float32)
        if cols.shape != (n_frames, 6):
            raise ValueError(f"frame_colors must be ({n_frames}, 6)")
    img = torch.empty((camera.num_rows, camera.num_cols, 3), dtype=torch.float32, device="cuda")
    errors = C.c_int64()
    s = stream if stream is not None else torch.cuda.current_stream()
    N.check(N.lib().mlgpu_render_pinhole(scene._h, world_frame, C.byref(camera.to_c()), C.byref(config.to_c()),
                                         cols.ctypes.data if cols is not None else None, n_frames, seed,
                                         C.c_void_p(img.data_ptr()), C.byref(errors), C.c_void_p(s.cuda_stream)),
            "mlgpu_render_pinhole")
    if errors.value:
        raise ValueError(f"{errors.value} pixel rays met a surface function that does not cover 533 nm")
    return img


def to_rgb8(image) -> np.ndarray:
    """The bytes of PortablePixMap's append_uint8 (PortablePixMap.cpp:10-28):
    each float converted to uint8 by truncation."""
    a = image.detach().cpu().numpy() if isinstance(image, torch.Tensor) else np.asarray(image, np.float32)
    return np.trunc(a).astype(np.int64).astype(np.uint8)


def write_ppm(image, path) -> None:
    """write_image_to_path (PortablePixMap.cpp:10-28,96-110): P6, 8 bit."""
    rgb = to_rgb8(image)
    rows, cols = rgb.shape[:2]
    head = f"P6\n# CREATOR: MERLICT\n{cols} {rows}\n255\n".encode()
    Path(path).write_bytes(head + np.ascontiguousarray(rgb).tobytes())


def read_ppm(path) -> np.ndarray:
    """read_image_from_file (PortablePixMap.cpp:30-94): (rows, cols, 3) float32."""
    raw = Path(path).read_bytes()
    lines, pos = [], 0
    for _ in range(4):
        end = raw.index(b"\n", pos)
        lines.append(raw[pos:end].decode())
        pos = end + 1
    if lines[0].strip() != "P6":
        raise ValueError(f"expected magic 'P6', got {lines[0]!r}")
    cols, rows = (int(v) for v in lines[2].split()[:2])
    if int(lines[3]) != 255:
        raise ValueError("expected colour depth 255")
    data = np.frombuffer(raw[pos:pos + rows * cols * 3], np.uint8)
    return data.reshape(rows, cols, 3).astype(np.float32)
