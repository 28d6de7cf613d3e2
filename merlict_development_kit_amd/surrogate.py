"""The learned surrogate of the propagator (mlgpu_surrogate): an MLP from a
photon's normalised (support x, y, direction cx, cy, wavelength) to its
survival logit and an 8-component Gaussian mixture over sensor-plane
position and arrival time, trained on the GPU by the library's own bf16
matrix-core kernels on photons the plain propagator produced."""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N

NUM_OUTPUTS = 57
NUM_COMPONENTS = 8


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream(stream) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Surrogate:
    def __init__(self, seed: int = 0):
        self._h = C.c_void_p()
        N.check(N.lib().mlgpu_surrogate_create(seed, C.byref(self._h)), "mlgpu_surrogate_create")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().mlgpu_surrogate_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @staticmethod
    def num_params() -> int:
        return int(N.lib().mlgpu_surrogate_num_params())

    def params(self, device="cuda") -> torch.Tensor:
        out = torch.empty(self.num_params(), dtype=torch.float32, device=device)
        N.check(N.lib().mlgpu_surrogate_get_params(self._h, _ptr(out), _stream(None)), "mlgpu_surrogate_get_params")
        torch.cuda.synchronize()
        return out

    def set_params(self, params: torch.Tensor) -> None:
        p = params.to(dtype=torch.float32).contiguous()
        if p.numel() != self.num_params():
            raise ValueError("wrong number of parameters")
        N.check(N.lib().mlgpu_surrogate_set_params(self._h, _ptr(p), _stream(None)), "mlgpu_surrogate_set_params")
        torch.cuda.synchronize()

    @staticmethod
    def _check(features: torch.Tensor, targets: torch.Tensor = None):
        if features.dtype != torch.float32 or features.dim() != 2 or features.shape[1] != 5 \
                or not features.is_contiguous() or features.device.type != "cuda":
            raise ValueError("features must be a contiguous (n, 5) float32 tensor on the GPU")
        if targets is not None and (targets.dtype != torch.float32 or tuple(targets.shape) != (len(features), 4)
                                    or not targets.is_contiguous() or targets.device != features.device):
            raise ValueError("targets must be a contiguous (n, 4) float32 tensor on the features' device")

    @property
    def inference(self) -> str:
        """Inference precision of forward and the hybrid: "bf16" (the
        default) or "fp8" (layers 1.. on the fp8 matrix instructions with
        e4m3 weights scaled per row by a power of two and e4m3 activations;
        layer 0 in bf16; fp32 accumulation)."""
        return "fp8" if N.lib().mlgpu_surrogate_get_inference(self._h) == 1 else "bf16"

    @inference.setter
    def inference(self, precision: str) -> None:
        codes = {"bf16": 0, "fp8": 1}
        if precision not in codes:
            raise ValueError(f"precision must be one of {sorted(codes)}")
        N.check(N.lib().mlgpu_surrogate_set_inference(self._h, codes[precision]), "mlgpu_surrogate_set_inference")

    def forward(self, features: torch.Tensor, stream=None) -> torch.Tensor:
        self._check(features)
        out = torch.empty((len(features), NUM_OUTPUTS), dtype=torch.float32, device=features.device)
        N.check(N.lib().mlgpu_surrogate_forward(self._h, _ptr(features), len(features), _ptr(out), _stream(stream)),
                "mlgpu_surrogate_forward")
        return out

    def gradients(self, features: torch.Tensor, targets: torch.Tensor, stream=None):
        """(mean loss, gradient of the flat parameters)."""
        self._check(features, targets)
        g = torch.empty(self.num_params(), dtype=torch.float32, device=features.device)
        loss = torch.empty(1, dtype=torch.float32, device=features.device)
        N.check(N.lib().mlgpu_surrogate_gradients(self._h, _ptr(features), _ptr(targets), len(features), _ptr(g),
                                                  _ptr(loss), _stream(stream)), "mlgpu_surrogate_gradients")
        return loss / len(features), g

    def set_survival_scale(self, scale: float) -> None:
        """Factor on the loss gradient wrt the survival logit in training
        (mlgpu_surrogate_set_survival_scale; the loss value is unchanged)."""
        N.check(N.lib().mlgpu_surrogate_set_survival_scale(self._h, float(scale)),
                "mlgpu_surrogate_set_survival_scale")

    def train_step(self, features: torch.Tensor, targets: torch.Tensor, learning_rate: float,
                   loss: torch.Tensor = None, stream=None) -> None:
        """One Adam step; `loss` (1-element float32 on the GPU) receives the
        summed loss of the batch before the step."""
        self._check(features, targets)
        N.check(N.lib().mlgpu_surrogate_train_step(self._h, _ptr(features), _ptr(targets), len(features),
                                                   float(learning_rate), _ptr(loss) if loss is not None else None,
                                                   _stream(stream)), "mlgpu_surrogate_train_step")

    def set_region(self, domain: "SurrogateDomain" = None, cells: "HexCells" = None) -> None:
        """Train with the mixture truncated to `cells`, the cells its draws
        are restricted to in the hybrid (mlgpu_surrogate_set_region); no
        arguments clear the region.  The cells' table is kept alive here."""
        if cells is None:
            N.check(N.lib().mlgpu_surrogate_set_region(self._h, None, None), "mlgpu_surrogate_set_region")
            self._region = None
            return
        N.check(N.lib().mlgpu_surrogate_set_region(self._h, C.byref(domain.c_struct()), C.byref(cells.c_struct())),
                "mlgpu_surrogate_set_region")
        self._region = cells


# ------------------------------------------------------------ domain, cells
import math  # noqa: E402
from dataclasses import dataclass, replace  # noqa: E402

import numpy as np  # noqa: E402

from .arrivals import ArrivalBinning  # noqa: E402
from .propagation import GpuScene, PhotonBatch, PropagationConfig, PropagationResult  # noqa: E402
from .scenery import SceneDescription  # noqa: E402
from . import sources as S  # noqa: E402


@dataclass(frozen=True)
class SurrogateDomain:
    """mlgpu_surrogate_domain: where the surrogate is trained and trusted,
    and the units of its targets."""
    support_radius: float
    max_angle: float
    lambda_min: float
    lambda_max: float
    image_scale: float = 0.0
    position_scale: float = 0.01
    time_offset: float = 0.0
    time_scale: float = 1e-9

    def c_struct(self) -> N.SurrogateDomain:
        return N.SurrogateDomain(self.support_radius, self.max_angle, self.lambda_min, self.lambda_max,
                                 self.image_scale, self.position_scale, self.time_offset, self.time_scale)

    def source(self, height: float) -> S.Source:
        """Photons spanning the domain: supports uniform on the disc at z =
        height, directions uniform on the cap within max_angle of -z, lambda
        with pdf ~ lambda^-2."""
        return S.Source("disc", "cap", self.support_radius, self.max_angle, self.lambda_min, self.lambda_max, 0.0,
                        S.FLIP, (0.0, 0.0, height))


def fact_domain() -> SurrogateDomain:
    """The FACT telescope's domain: supports within 1.1 x the 2.1 m
    mirror radius (fact_propagation.cpp:295), directions within 1.2 x the
    camera's 2.2 deg field radius, lambda in [280, 700] nm; positions
    relative to the nominal image point in units of 1 cm.  The image scale
    and the time units are set from data by :func:`train`."""
    return SurrogateDomain(1.1 * 2.1, math.radians(1.2 * 2.2), 280e-9, 700e-9)


class HexCells:
    """mlgpu_hex_cells of a scene's sensors.  Their world centres (frame
    poses) must lie near the sites of a hexagonal lattice (within a fifth of
    a lattice step; the FACT camera's pixel table deviates by up to 0.4 mm
    from a 9.5 mm lattice), one sensor per site; the cells are centred on
    the sensors.  mode "hexagon": a cell is the hexagon a photon must enter
    to reach its sensor, the entry hexagon of the light guide in front of a
    pixel (a HexPlane sibling without a sensor id) or the bare pixel's own
    HexPlane, and a photon in a gap between them reaches none.  mode
    "nearest" (the default): a cell holds the points nearest to its centre
    within a hexagon just wider than half the widest spacing of neighbouring
    sensors, so the cells close every gap inside the camera (a photon that
    falls between two light-guide entries meets a side wall and is refracted
    into a guide) and end about half a pitch beyond its rim.  Sensors that
    are not hexagons of one size get the cells of their lattice sites."""

    def __init__(self, desc: SceneDescription, device="cuda", mode: str = "nearest"):
        if mode not in ("hexagon", "nearest"):
            raise ValueError("mode must be 'hexagon' or 'nearest'")
        frames = desc.sensor_frames()
        poses = desc.frame_poses()
        centres = poses[frames][:, :2, 3]
        d = centres[:, None, :] - centres[None, :, :]
        dist = np.hypot(d[..., 0], d[..., 1])
        np.fill_diagonal(dist, np.inf)
        pitch = float(np.median(dist.min(axis=1)))
        # first lattice vector: the nearest-neighbour offset of the most
        # central cell at the smallest angle
        c0 = int(np.argmin(np.hypot(centres[:, 0], centres[:, 1])))
        nb = np.nonzero(np.abs(dist[c0] - pitch) < 0.05 * pitch)[0]
        vecs = centres[nb] - centres[c0]
        ang = np.mod(np.arctan2(vecs[:, 1], vecs[:, 0]), 2 * np.pi)
        a1 = vecs[int(np.argmin(ang))]
        rot = np.array([[0.5, -math.sqrt(3) / 2], [math.sqrt(3) / 2, 0.5]])
        a2 = rot @ a1
        origin = centres[c0]
        A = np.stack([a1, a2], axis=1)
        ij = np.linalg.solve(A, (centres - origin).T).T
        rij = np.rint(ij)
        if np.abs(ij - rij).max() > 0.2:
            raise ValueError("sensor centres are not near the sites of a hexagonal lattice")
        radius = int(np.abs(rij).max())
        table = np.full((2 * radius + 1) ** 2, -1, dtype=np.int32)
        idx = ((rij[:, 0] + radius) * (2 * radius + 1) + (rij[:, 1] + radius)).astype(np.int64)
        if len(np.unique(idx)) != len(idx):
            raise ValueError("two sensors share a lattice site")
        table[idx] = np.arange(len(frames), dtype=np.int32)
        self.a1, self.a2, self.origin, self.radius = a1, a2, origin, radius
        self.table = torch.tensor(table, device=device)
        self.pitch = pitch
        r_in = self._hexagon_inner_radius(desc, frames)
        self.mode = mode if r_in > 0 else "lattice"
        if self.mode == "nearest":
            # half the widest separation of neighbouring centres along the
            # hexagons' flat-side normals, plus 0.1 %
            nrm = np.stack([a1, rot @ a1, rot @ rot @ a1]) / np.linalg.norm(a1)
            close = (dist < 1.5 * pitch) & np.isfinite(dist)
            sep = np.abs(np.einsum("ijk,lk->ijl", d, nrm)).max(axis=2)
            r_in = 0.5 * float(sep[close].max()) * 1.001
        self.inner_radius = r_in if self.mode != "lattice" else 0.0
        self.centres = (torch.tensor(centres.astype(np.float32).reshape(-1), device=device)
                        if self.mode != "lattice" else None)

    @staticmethod
    def _hexagon_inner_radius(desc: SceneDescription, frames: np.ndarray) -> float:
        """Inner radius of the hexagon a photon enters to reach a sensor, the
        same for every sensor (light-guide entries, else bare HexPlane
        pixels), or 0."""
        a = desc.arrays()
        kind, mother, sensor, params = a["kind"], a["mother"], a["sensor_id"], a["params"]
        hexplane = N.KIND["HexPlane"]
        entry = (kind == hexplane) & (sensor < 0) & (mother >= 0)
        count = np.zeros(len(kind), np.int64)
        np.add.at(count, mother[entry], 1)
        if np.all(count[mother[frames]] == 1):
            first = np.full(len(kind), -1, np.int64)
            first[mother[entry]] = np.nonzero(entry)[0]
            outer = params[first[mother[frames]], 0]
        elif np.all(kind[frames] == hexplane):
            outer = params[frames, 0]
        else:
            return 0.0
        return float(outer[0] * math.sqrt(3) / 2) if np.all(outer == outer[0]) else 0.0

    def c_struct(self) -> N.HexCells:
        c = N.HexCells()
        for k in range(2):
            c.a1[k], c.a2[k], c.origin[k] = float(self.a1[k]), float(self.a2[k]), float(self.origin[k])
        c.radius = self.radius
        c.cell = self.table.data_ptr()
        c.centre = self.centres.data_ptr() if self.centres is not None else None
        c.inner_radius = self.inner_radius
        return c


def prepare(domain: SurrogateDomain, photons: PhotonBatch, result: PropagationResult = None,
            poses: torch.Tensor = None, focal: PropagationResult = None, focal_poses: torch.Tensor = None,
            stream=None):
    """(features, targets) of photons and their propagation results; with
    `focal` (the same photons and config through the focal-plane scene,
    frame poses `focal_poses`) the position targets are the focal-plane
    crossings (mlgpu_surrogate_prepare)."""
    n = len(photons)
    dev = photons.device
    feats = torch.empty((n, 5), dtype=torch.float32, device=dev)
    targs = torch.empty((n, 4), dtype=torch.float32, device=dev) if result is not None else None
    ps = photons.c_struct()
    rs = result.c_struct() if result is not None else None
    fs = focal.c_struct() if focal is not None else None
    N.check(N.lib().mlgpu_surrogate_prepare(C.byref(domain.c_struct()), C.byref(ps),
                                            C.byref(rs) if rs is not None else None,
                                            _ptr(poses) if poses is not None else None,
                                            C.byref(fs) if fs is not None else None,
                                            _ptr(focal_poses) if focal_poses is not None else None, n, _ptr(feats),
                                            _ptr(targs) if targs is not None else None, _stream(stream)),
            "mlgpu_surrogate_prepare")
    return feats, targs


def train(scene: GpuScene, desc: SceneDescription, surrogate: Surrogate, domain: SurrogateDomain, n_photons: int,
          height: float = 10.0, batch: int = 16384, learning_rate: float = 1e-3, final_learning_rate: float = 3e-5,
          seed: int = 0, chunk: int = 1 << 21, max_interactions: int = 5, device="cuda", log_every: int = 0,
          cells: "HexCells" = None, focal_desc: SceneDescription = None, survival_scale: float = 10.0):
    """Train on n_photons drawn from the domain and propagated by the plain
    propagator (Philox seed `seed`, ids 0 .. n_photons - 1), Adam with a
    cosine learning-rate decay.  With `cells`, lost photons are explained by
    the mixture's mass off the cells as well (Surrogate.set_region), as the
    hybrid's draws treat them.  With `focal_desc` (the scene with a sensor at
    the focal plane, e.g. SceneDescription.fact(focal_plane=True)) the same
    photons are propagated through it too and the position targets are
    their focal-plane crossings.  The gradient wrt the survival logit is
    scaled by `survival_scale` (Surrogate.set_survival_scale): unscaled, the
    shared layers barely learn the survival's few-per-cent dependence on the
    wavelength (the mirror's reflectivity) next to the position terms.
    Returns the domain (its time offset and scale set from the data when
    they were unset) and the loss history (mean loss per batch, every
    log_every steps)."""
    poses = torch.tensor(desc.frame_poses().reshape(-1), dtype=torch.float64, device=device)
    src = domain.source(height)
    cfg = PropagationConfig(max_interactions, seed=seed)
    focal_scene = GpuScene(focal_desc) if focal_desc is not None else None
    focal_poses = (torch.tensor(focal_desc.frame_poses().reshape(-1), dtype=torch.float64, device=device)
                   if focal_desc is not None else None)

    def targets(dom, ph, c):
        res = scene.propagate(ph, c)
        foc = focal_scene.propagate(ph, c) if focal_scene is not None else None
        return res, prepare(dom, ph, res, poses, foc, focal_poses)

    if domain.time_offset == 0.0 or domain.image_scale == 0.0:
        # time units and the image scale from a first sample: the least-squares
        # slope of the sensor-plane position against (dx / dz, dy / dz)
        ph = src.generate(1 << 18, seed=seed ^ 0x5EED, device=device)
        res, (f, t) = targets(replace(domain, image_scale=0.0, position_scale=1.0), ph, cfg)
        det = res.sensor_id >= 0
        tof = res.time_of_flight[det]
        sa = math.sin(domain.max_angle)
        d = f[det][:, 2:4].double() * sa
        dz = -torch.sqrt(torch.clamp(1.0 - (d * d).sum(dim=1), min=1e-12))
        u = (d / dz[:, None]).reshape(-1)
        p = t[det][:, 1:3].double().reshape(-1)
        scale = float((u * p).sum() / (u * u).sum())
        domain = replace(domain, image_scale=scale, time_offset=float(tof.mean()),
                         time_scale=float(max(tof.std().item(), 1e-12)))
    if cells is not None:
        surrogate.set_region(domain, cells)
    surrogate.set_survival_scale(survival_scale)
    steps = max(1, n_photons // batch)
    step = 0
    history = []
    loss = torch.zeros(1, dtype=torch.float32, device=device)
    for c0 in range(0, n_photons, chunk):
        m = min(chunk, n_photons - c0)
        ph = src.generate(m, seed=seed, photon_id_offset=c0, device=device)
        _, (f, t) = targets(domain, ph, PropagationConfig(max_interactions, seed=seed, photon_id_offset=c0))
        for b0 in range(0, m - batch + 1, batch):
            lr = final_learning_rate + 0.5 * (learning_rate - final_learning_rate) * (1 + math.cos(math.pi * step / steps))
            surrogate.train_step(f[b0:b0 + batch], t[b0:b0 + batch], lr, loss)
            if log_every and step % log_every == 0:
                history.append((step, loss.item() / batch))
            step += 1
    return domain, history


def propagate_hybrid(scene: GpuScene, surrogate: Surrogate, domain: SurrogateDomain, cells: HexCells,
                     photons: PhotonBatch, binning: ArrivalBinning, config: PropagationConfig = PropagationConfig(),
                     max_entropy: float = 0.0, seed: int = 0, counts: torch.Tensor = None, world_frame: int = 0,
                     stream=None):
    """Histogram of the photons' arrivals, each drawn from the surrogate or,
    out of its domain or above the entropy bound, propagated.  Returns
    (counts, number of propagated photons)."""
    if counts is None:
        counts = binning.empty(photons.device)
    ps = photons.c_struct()
    cfg = config.c_struct()
    hc = N.HybridConfig(max_entropy, seed)
    h = binning.c_struct(counts)
    nf = C.c_int64()
    N.check(N.lib().mlgpu_propagate_hybrid(scene.handle, surrogate.handle, C.byref(domain.c_struct()),
                                           C.byref(cells.c_struct()), world_frame, C.byref(ps), len(photons),
                                           C.byref(cfg), C.byref(hc), C.byref(h), C.byref(nf), _stream(stream)),
            "mlgpu_propagate_hybrid")
    return counts, nf.value
