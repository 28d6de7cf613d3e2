"""Python access to the oracle and to the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py, never by the product.

* :class:`OracleScene` -- ctypes binding of ``liboracle.so`` (oracle.c), the
  FP64 restatement of merlict's propagation.
* :func:`ref_propagate` & co. -- run ``_ref/merlict_golden``, the harness
  compiled against the reference's own sources, on files.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import json
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_BIN = HERE / "_ref" / "merlict_golden"

_P_I32 = C.POINTER(C.c_int32)
_P_F64 = C.POINTER(C.c_double)
_P_U8 = C.POINTER(C.c_uint8)


class _Results(C.Structure):
    _fields_ = [
        ("final_frame", _P_I32), ("sensor_id", _P_I32), ("final_interaction", _P_U8),
        ("num_interactions", _P_U8), ("error", _P_U8), ("time_of_flight", _P_F64),
        ("x", _P_F64), ("y", _P_F64), ("cx", _P_F64), ("cy", _P_F64), ("accumulated_distance", _P_F64),
    ]


class _Photons(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "support_x", "support_y", "support_z", "direction_x", "direction_y", "direction_z", "wavelength")]


class Source(C.Structure):
    """mlgpu_source of include/mlgpu.h."""
    _fields_ = [
        ("support", C.c_int32), ("direction", C.c_int32), ("radius", C.c_double), ("angle", C.c_double),
        ("lambda_min", C.c_double), ("lambda_max", C.c_double), ("start_distance", C.c_double),
        ("rot", C.c_double * 9), ("pos", C.c_double * 3), ("angle_width", C.c_double), ("spectrum", C.c_int32),
        ("pad", C.c_int32),
    ]


class Histogram(C.Structure):
    """mlgpu_histogram of include/mlgpu.h."""
    _fields_ = [
        ("counts", C.c_void_p), ("num_sensors", C.c_int32), ("num_bins", C.c_int32), ("t_start", C.c_double),
        ("bin_width", C.c_double), ("jitter_sigma", C.c_double), ("seed", C.c_uint64),
        ("photon_id_offset", C.c_uint64),
    ]


PHOTON_FIELDS = ("support_x", "support_y", "support_z", "direction_x", "direction_y", "direction_z", "wavelength")

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing; run `make -C oracle oracle`")
        L = C.CDLL(str(LIB_PATH))
        L.oracle_scene_create.restype = C.c_void_p
        L.oracle_scene_create.argtypes = [C.c_void_p]
        L.oracle_scene_destroy.argtypes = [C.c_void_p]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_num_frames.argtypes = [C.c_void_p]
        L.oracle_tree_dump.argtypes = [C.c_void_p, _P_I32, _P_I32, _P_F64, _P_F64, C.c_int]
        L.oracle_first_hit.argtypes = [C.c_void_p, C.c_int32, _P_F64, _P_F64, _P_I32, _P_F64, _P_F64,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.oracle_propagate.argtypes = [C.c_void_p, C.c_int32, C.POINTER(_Photons), C.c_int64, C.c_uint32,
                                       C.c_uint64, C.c_uint64, C.POINTER(_Results), C.c_int,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.oracle_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.oracle_uniform.restype = C.c_double
        L.oracle_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32]
        L.oracle_func_evaluate.argtypes = [C.c_void_p, C.c_int32, C.c_double, _P_F64]
        L.oracle_generate_photons.argtypes = [C.POINTER(Source), C.c_int64, C.c_uint64, C.c_uint64,
                                              C.POINTER(_P_F64)]
        L.oracle_histogram_arrivals.argtypes = [_P_I32, _P_U8, _P_F64, C.c_int64, C.POINTER(Histogram),
                                                C.POINTER(C.c_uint32)]
        L.oracle_arrival_jitter.restype = C.c_double
        L.oracle_arrival_jitter.argtypes = [C.c_uint64, C.c_uint64]
        L.oracle_corsika_photons.restype = C.c_int64
        L.oracle_corsika_photons.argtypes = [C.POINTER(C.c_float), C.c_int64, C.c_uint64, C.c_uint64,
                                             C.POINTER(_P_F64), _P_I32, C.c_int64]
        L.oracle_log_series.restype = C.c_double
        L.oracle_log_series.argtypes = [C.c_double]
        L.oracle_sincos_turn.argtypes = [C.c_double, _P_F64, _P_F64]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def photons_np(d: dict) -> dict:
    return {k: np.ascontiguousarray(d[k], dtype=np.float64) for k in PHOTON_FIELDS}


class OracleScene:
    """The oracle's scene, built from a product SceneDescription's C view."""

    def __init__(self, desc):
        self.desc = desc  # keeps the arrays alive
        self._h = lib().oracle_scene_create(C.addressof(desc.view))
        if not self._h:
            raise RuntimeError("oracle_scene_create: " + lib().oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().oracle_scene_destroy(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def num_frames(self) -> int:
        return lib().oracle_num_frames(self._h)

    def render_pinhole(self, camera, config, frame_colors=None, seed: int = 0, world: int = 0):
        """(image rows x cols x 3 float32, errors) of oracle_render_pinhole;
        camera / config: ctypes structs laid out as mlgpu_pinhole_camera /
        mlgpu_render_config."""
        img = np.zeros((camera.num_rows, camera.num_cols, 3), np.float32)
        cols = None if frame_colors is None else np.ascontiguousarray(frame_colors, np.float32)
        L = lib()
        L.oracle_render_pinhole.restype = C.c_int64
        L.oracle_render_pinhole.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                            C.c_void_p]
        k = L.oracle_render_pinhole(self._h, world, C.addressof(camera), C.addressof(config),
                                    cols.ctypes.data if cols is not None else None, seed, img.ctypes.data)
        if k < 0:
            raise ValueError("bad render arguments")
        return img, int(k)

    def tree_dump(self) -> dict:
        n = self.num_frames
        frame = np.zeros(n, np.int32)
        nch = np.zeros(n, np.int32)
        t2w = np.zeros(12 * n, np.float64)
        rad = np.zeros(n, np.float64)
        k = lib().oracle_tree_dump(self._h, _ptr(frame, _P_I32), _ptr(nch, _P_I32), _ptr(t2w, _P_F64),
                                   _ptr(rad, _P_F64), n)
        assert k == n
        return {"frame": frame, "n_children": nch, "frame2world": t2w.reshape(n, 12), "radius": rad}

    def first_hit(self, rays: dict, root_frame: int = 0) -> dict:
        p = photons_np(rays)
        n = len(p["support_x"])
        out = {"frame": np.zeros(n, np.int32), "distance": np.zeros(n), "position": np.zeros((n, 3))}
        st, lt = C.c_int64(0), C.c_int64(0)
        for i in range(n):
            s = np.array([p["support_x"][i], p["support_y"][i], p["support_z"][i]])
            d = np.array([p["direction_x"][i], p["direction_y"][i], p["direction_z"][i]])
            f, dist, pos = C.c_int32(), C.c_double(), np.zeros(3)
            rc = lib().oracle_first_hit(self._h, root_frame, _ptr(s, _P_F64), _ptr(d, _P_F64), C.byref(f),
                                        C.byref(dist), _ptr(pos, _P_F64), C.byref(st), C.byref(lt))
            if rc:
                raise RuntimeError(lib().oracle_last_error().decode())
            out["frame"][i] = f.value
            out["distance"][i] = dist.value
            out["position"][i] = pos
        out["sphere_tests"], out["leaf_tests"] = st.value, lt.value
        w = _Work()
        lib().oracle_last_work(C.byref(w))
        out["work"] = {name: int(getattr(w, name)) for name, _ in _Work._fields_}
        return out

    def propagate(self, photons: dict, max_num_interactions: int = 5, seed: int = 0, photon_id_offset: int = 0,
                  threads: int = 1, world_frame: int = 0) -> dict:
        p = photons_np(photons)
        n = len(p["support_x"])
        ph = _Photons(*(p[k].ctypes.data for k in PHOTON_FIELDS))
        out = {
            "final_frame": np.zeros(n, np.int32), "sensor_id": np.zeros(n, np.int32),
            "final_interaction": np.zeros(n, np.uint8), "num_interactions": np.zeros(n, np.uint8),
            "error": np.zeros(n, np.uint8), "time_of_flight": np.zeros(n), "x": np.zeros(n), "y": np.zeros(n),
            "cx": np.zeros(n), "cy": np.zeros(n), "accumulated_distance": np.zeros(n),
        }
        r = _Results(_ptr(out["final_frame"], _P_I32), _ptr(out["sensor_id"], _P_I32),
                     _ptr(out["final_interaction"], _P_U8), _ptr(out["num_interactions"], _P_U8),
                     _ptr(out["error"], _P_U8), _ptr(out["time_of_flight"], _P_F64), _ptr(out["x"], _P_F64),
                     _ptr(out["y"], _P_F64), _ptr(out["cx"], _P_F64), _ptr(out["cy"], _P_F64),
                     _ptr(out["accumulated_distance"], _P_F64))
        st, lt = C.c_int64(0), C.c_int64(0)
        rc = lib().oracle_propagate(self._h, world_frame, C.byref(ph), n, max_num_interactions, seed,
                                    photon_id_offset, C.byref(r), threads, C.byref(st), C.byref(lt))
        if rc:
            raise RuntimeError(lib().oracle_last_error().decode())
        out["sphere_tests"], out["leaf_tests"] = st.value, lt.value
        w = _Work()
        lib().oracle_last_work(C.byref(w))
        out["work"] = {name: int(getattr(w, name)) for name, _ in _Work._fields_}
        return out

    def func_evaluate(self, f: int, x: float):
        y = C.c_double()
        rc = lib().oracle_func_evaluate(self._h, f, x, C.byref(y))
        return (None if rc else y.value)


class _Work(C.Structure):
    """oracle_work: event counts of the reference's algorithm (oracle.h)"""
    _fields_ = [(name, C.c_int64) for name in (
        "queries", "sphere_tests", "leaf_cap", "leaf_planar", "leaf_cylinder", "leaf_sphere", "func_lookups",
        "func_compares", "fresnel", "bounces")]


def generate_photons(src: Source, n: int, seed: int = 0, photon_id_offset: int = 0) -> dict:
    """Photons of a source, restated on the CPU (oracle_generate_photons)."""
    out = {k: np.empty(n, dtype=np.float64) for k in PHOTON_FIELDS}
    arr = (_P_F64 * 7)(*[_ptr(out[k], _P_F64) for k in PHOTON_FIELDS])
    lib().oracle_generate_photons(C.byref(src), n, seed, photon_id_offset, arr)
    return out


def histogram_arrivals(results: dict, num_sensors: int, num_bins: int, t_start: float, bin_width: float,
                       jitter_sigma: float = 0.0, seed: int = 0, photon_id_offset: int = 0) -> np.ndarray:
    """Arrival-time histogram (num_sensors, num_bins) of oracle results."""
    sensor = np.ascontiguousarray(results["sensor_id"], dtype=np.int32)
    err = np.ascontiguousarray(results["error"], dtype=np.uint8)
    tof = np.ascontiguousarray(results["time_of_flight"], dtype=np.float64)
    counts = np.zeros(num_sensors * num_bins, dtype=np.uint32)
    h = Histogram(None, num_sensors, num_bins, t_start, bin_width, jitter_sigma, seed, photon_id_offset)
    lib().oracle_histogram_arrivals(_ptr(sensor, _P_I32), _ptr(err, _P_U8), _ptr(tof, _P_F64), len(sensor),
                                    C.byref(h), counts.ctypes.data_as(C.POINTER(C.c_uint32)))
    return counts.reshape(num_sensors, num_bins)


def corsika_photons(bunches: np.ndarray, seed: int = 0, id_offset: int = 0):
    """Photons of CORSIKA bunches (n x 8 float32), restated on the CPU
    (oracle_corsika_photons): (photons dict, truth ids).  Raises ValueError
    for a weight outside [0, 16]."""
    b = np.ascontiguousarray(bunches, dtype=np.float32)
    n = len(b)
    cap = int(np.floor(np.clip(b[:, 6].astype(np.float64), 0, 16)).sum()) + n if n else 0
    out = {k: np.empty(cap, dtype=np.float64) for k in PHOTON_FIELDS}
    truth = np.empty(cap, dtype=np.int32)
    arr = (_P_F64 * 7)(*[_ptr(out[k], _P_F64) for k in PHOTON_FIELDS])
    m = lib().oracle_corsika_photons(b.ctypes.data_as(C.POINTER(C.c_float)), n, seed, id_offset, arr,
                                     _ptr(truth, _P_I32), cap)
    if m == -1:
        raise ValueError("CORSIKA bunch weight outside [0, 16]")
    if m < 0:
        raise RuntimeError("capacity")
    return {k: v[:m] for k, v in out.items()}, truth[:m]


def ref_corsika_extract(path) -> np.ndarray:
    """Every photon bunch of a CORSIKA eventio file, read by the reference."""
    with tempfile.TemporaryDirectory() as d:
        o = Path(d) / "b.bin"
        _run_ref(["corsika-extract", path, o])
        raw = o.read_bytes()
    n = int(np.frombuffer(raw[:8], np.int64)[0])
    return np.frombuffer(raw[8:], np.float32).reshape(n, 8).copy()


def ref_corsika(bunches: np.ndarray, seed: int = 0, id_offset: int = 0) -> dict:
    """The reference's EventIoPhotonFactory on each bunch (Philox stream 5 of
    (seed, id_offset + bunch)): status (1 = bad weight), count, photons, truth."""
    b = np.ascontiguousarray(bunches, dtype=np.float32)
    with tempfile.TemporaryDirectory() as d:
        i, o = Path(d) / "b.bin", Path(d) / "o.bin"
        i.write_bytes(np.array([len(b)], np.int64).tobytes() + b.tobytes())
        _run_ref(["corsika", i, seed, id_offset, o])
        raw = o.read_bytes()
    n = int(np.frombuffer(raw[:8], np.int64)[0])
    pos = 8
    status = np.frombuffer(raw[pos:pos + n], np.uint8).copy(); pos += n
    count = np.frombuffer(raw[pos:pos + 4 * n], np.uint32).copy(); pos += 4 * n
    m = int(np.frombuffer(raw[pos:pos + 8], np.int64)[0]); pos += 8
    ph = np.frombuffer(raw[pos:pos + 56 * m], np.float64).reshape(m, 7); pos += 56 * m
    truth = np.frombuffer(raw[pos:pos + 4 * m], np.int32).copy()
    return {"status": status, "count": count, "photons": {k: ph[:, j].copy() for j, k in enumerate(PHOTON_FIELDS)},
            "truth": truth}


def arrival_jitter(seed: int, photon_id: int) -> float:
    return lib().oracle_arrival_jitter(seed, photon_id)


def log_series(x: float) -> float:
    return lib().oracle_log_series(x)


def sincos_turn(u: float) -> tuple:
    s, c = C.c_double(), C.c_double()
    lib().oracle_sincos_turn(u, C.byref(s), C.byref(c))
    return s.value, c.value


def philox4x32_10(counter, key):
    c = (C.c_uint32 * 4)(*counter)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return tuple(o)


def uniform(seed: int, photon_id: int, draw: int, stream: int = 0) -> float:
    return lib().oracle_uniform(seed, stream, photon_id, draw)


# --------------------------------------------------------------- file I/O
def write_photons(path, photons: dict) -> None:
    p = photons_np(photons)
    n = len(p["support_x"])
    with open(path, "wb") as f:
        f.write(b"MLGPUPH1")
        f.write(np.int64(n).tobytes())
        for k in PHOTON_FIELDS:
            f.write(p[k].tobytes())


def read_photons(path) -> dict:
    with open(path, "rb") as f:
        assert f.read(8) == b"MLGPUPH1"
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        return {k: np.frombuffer(f.read(8 * n), np.float64).copy() for k in PHOTON_FIELDS}


def read_results(path) -> dict:
    with open(path, "rb") as f:
        magic = f.read(8)
        assert magic in (b"MLGPURS1", b"MLGPURS2")
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        out = {}
        keys = [("final_frame", np.int32), ("sensor_id", np.int32), ("final_interaction", np.uint8),
                ("num_interactions", np.uint8), ("error", np.uint8), ("time_of_flight", np.float64),
                ("x", np.float64), ("y", np.float64), ("cx", np.float64), ("cy", np.float64)]
        if magic == b"MLGPURS2":
            keys.append(("accumulated_distance", np.float64))
        for k, t in keys:
            out[k] = np.frombuffer(f.read(np.dtype(t).itemsize * n), t).copy()
        return out


def write_results(path, r: dict) -> None:
    n = len(r["final_frame"])
    with open(path, "wb") as f:
        f.write(b"MLGPURS1")
        f.write(np.int64(n).tobytes())
        for k, t in (("final_frame", np.int32), ("sensor_id", np.int32), ("final_interaction", np.uint8),
                     ("num_interactions", np.uint8), ("error", np.uint8), ("time_of_flight", np.float64),
                     ("x", np.float64), ("y", np.float64), ("cx", np.float64), ("cy", np.float64)):
            f.write(np.ascontiguousarray(r[k], dtype=t).tobytes())


# ------------------------------------------------------- reference harness
def ref_available() -> bool:
    return REF_BIN.exists() and os.access(REF_BIN, os.X_OK)


def _run_ref(args, capture=False):
    proc = subprocess.run([str(REF_BIN), *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                          text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"merlict_golden {args[0]} failed: {proc.stderr.strip()}")
    return proc.stdout


def ref_propagate(desc, photons: dict, seed=0, max_num_interactions=5, photon_id_offset=0, threads=1) -> dict:
    with tempfile.TemporaryDirectory() as d:
        sp, pp, rp = Path(d) / "s.bin", Path(d) / "p.bin", Path(d) / "r.bin"
        desc.save(str(sp))
        write_photons(pp, photons)
        _run_ref(["propagate", sp, pp, seed, max_num_interactions, photon_id_offset, rp, threads])
        return read_results(rp)


def ref_tree(desc) -> dict:
    with tempfile.TemporaryDirectory() as d:
        sp, tp = Path(d) / "s.bin", Path(d) / "t.bin"
        desc.save(str(sp))
        _run_ref(["tree", sp, tp])
        with open(tp, "rb") as f:
            assert f.read(8) == b"MLGPUTR1"
            n = int(np.frombuffer(f.read(8), np.int64)[0])
            frame = np.frombuffer(f.read(4 * n), np.int32).copy()
            nch = np.frombuffer(f.read(4 * n), np.int32).copy()
            t2w = np.frombuffer(f.read(8 * 12 * n), np.float64).copy().reshape(n, 12)
            rad = np.frombuffer(f.read(8 * n), np.float64).copy()
        return {"frame": frame, "n_children": nch, "frame2world": t2w, "radius": rad}


def ref_first_hit(desc, rays: dict) -> dict:
    with tempfile.TemporaryDirectory() as d:
        sp, pp, hp = Path(d) / "s.bin", Path(d) / "p.bin", Path(d) / "h.bin"
        desc.save(str(sp))
        write_photons(pp, rays)
        _run_ref(["firsthit", sp, pp, hp])
        with open(hp, "rb") as f:
            assert f.read(8) == b"MLGPUFH1"
            n = int(np.frombuffer(f.read(8), np.int64)[0])
            frame = np.frombuffer(f.read(4 * n), np.int32).copy()
            dist = np.frombuffer(f.read(8 * n), np.float64).copy()
            pos = np.stack([np.frombuffer(f.read(8 * n), np.float64).copy() for _ in range(3)], axis=1)
        return {"frame": frame, "distance": dist, "position": pos}


def ref_reflector(focal_length, dc, max_outer, min_inner, hex_shape, facet_inner, gap) -> np.ndarray:
    out = _run_ref(["reflector", focal_length, dc, max_outer, min_inner, hex_shape, facet_inner, gap])
    return np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])


def ref_lens_maker(f, r, n) -> float:
    return float(_run_ref(["lensmaker", f, r, n]).strip())


def ref_lightfield(f, aperture, fov_deg, pixel_fov_deg, n_paxel, lens_n, bin_refl, housing) -> tuple:
    out = _run_ref(["lightfield", f, aperture, fov_deg, pixel_fov_deg, n_paxel, lens_n, bin_refl, housing])
    lines = out.strip().splitlines()
    head = lines[0].split()
    meta = {head[i]: float(head[i + 1]) for i in range(1, len(head) - 1, 2)}
    rows = np.array([[float(v) for v in line.split()] for line in lines[1:]])
    return meta, rows


def ref_bench(scene_path, photons_path, threads, mode, seed=0, results_path=None) -> str:
    args = ["bench", scene_path, photons_path, threads, mode, seed]
    if results_path is not None:
        args.append(results_path)
    return _run_ref(args)


# ------------------------------------------------------------ signal chain
def _sig_lib():
    L = lib()
    if not getattr(L, "_sig_ready", False):
        L.oracle_sensor_pipelines.restype = C.c_int64
        L.oracle_sensor_pipelines.argtypes = [_P_I32, _P_U8, _P_F64, _P_F64, _P_I32, C.c_int64, C.c_int32,
                                              C.POINTER(C.c_int64), _P_F64, _P_F64, _P_I32]
        L.oracle_photoelectric.restype = C.c_int64
        L.oracle_photoelectric.argtypes = [C.POINTER(C.c_int64), _P_F64, _P_F64, _P_I32, C.c_int32, _P_F64, _P_F64,
                                           C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                           C.POINTER(C.c_int64), _P_F64, _P_I32, C.c_int64]
        L.oracle_extract_pulses.argtypes = [C.POINTER(C.c_int64), _P_F64, C.c_int32, C.c_double, C.c_double,
                                            C.c_uint64, C.POINTER(C.c_int16)]
        L.oracle_inject_nsb.restype = C.c_int64
        L.oracle_inject_nsb.argtypes = [C.POINTER(C.c_int64), _P_F64, _P_F64, _P_I32, C.c_int32, C.POINTER(C.c_float),
                                        _P_F64, _P_F64, C.c_int32, C.c_double, C.c_double, C.c_uint64,
                                        C.POINTER(C.c_int64), _P_F64, _P_F64, _P_I32, C.c_int64,
                                        C.POINTER(C.c_double)]
        L.oracle_nsb_light_rate.restype = C.c_double
        L.oracle_nsb_light_rate.argtypes = [_P_F64, _P_F64, C.c_int32, C.c_double, C.c_double]
        L._sig_ready = True
    return L


def sensor_pipelines(sensor_id, error, time_of_flight, wavelength, truth_id, num_sensors):
    """(offsets, arrival_time, wavelength, truth_id) of the per-sensor
    pipelines (oracle_sensor_pipelines)."""
    s = np.ascontiguousarray(sensor_id, np.int32)
    e = np.ascontiguousarray(error, np.uint8)
    t = np.ascontiguousarray(time_of_flight, np.float64)
    lam = np.ascontiguousarray(wavelength, np.float64)
    tr = np.ascontiguousarray(truth_id, np.int32) if truth_id is not None else None
    n = len(s)
    off = np.zeros(num_sensors + 1, np.int64)
    ot, ol, otr = np.empty(n), np.empty(n), np.empty(n, np.int32)
    m = _sig_lib().oracle_sensor_pipelines(_ptr(s, _P_I32), _ptr(e, _P_U8), _ptr(t, _P_F64), _ptr(lam, _P_F64),
                                           _ptr(tr, _P_I32) if tr is not None else None, n, num_sensors,
                                           off.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(ot, _P_F64),
                                           _ptr(ol, _P_F64), _ptr(otr, _P_I32))
    return off, ot[:m], ol[:m], otr[:m]


def photoelectric(offsets, arrival_time, wavelength, truth_id, num_sensors, qe_x, qe_y, dark_rate, crosstalk,
                  exposure_time, seed, capacity=None):
    """(offsets, time, truth) of the converter's pulses (oracle_photoelectric)."""
    off = np.ascontiguousarray(offsets, np.int64)
    t = np.ascontiguousarray(arrival_time, np.float64)
    lam = np.ascontiguousarray(wavelength, np.float64)
    tr = np.ascontiguousarray(truth_id, np.int32)
    qx, qy = np.ascontiguousarray(qe_x, np.float64), np.ascontiguousarray(qe_y, np.float64)
    cap = capacity or (4 * len(t) + 64 * num_sensors + int(dark_rate * exposure_time * num_sensors * 4) + 64)
    oo = np.zeros(num_sensors + 1, np.int64)
    ot, otr = np.empty(cap), np.empty(cap, np.int32)
    k = _sig_lib().oracle_photoelectric(off.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(t, _P_F64),
                                        _ptr(lam, _P_F64), _ptr(tr, _P_I32), num_sensors, _ptr(qx, _P_F64),
                                        _ptr(qy, _P_F64), len(qx), dark_rate, crosstalk, exposure_time, seed,
                                        oo.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(ot, _P_F64),
                                        _ptr(otr, _P_I32), cap)
    if k == -1:
        raise IndexError("wavelength outside the quantum efficiency table")
    if k < 0:
        raise RuntimeError("capacity")
    return oo, ot[:k], otr[:k]


def extract_pulses(offsets, arrival_time, num_sensors, duration, std_dev, seed):
    """Time slice per pulse (-1: dropped) (oracle_extract_pulses)."""
    off = np.ascontiguousarray(offsets, np.int64)
    t = np.ascontiguousarray(arrival_time, np.float64)
    out = np.empty(len(t), np.int16)
    _sig_lib().oracle_extract_pulses(off.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(t, _P_F64), num_sensors,
                                     duration, std_dev, seed, out.ctypes.data_as(C.POINTER(C.c_int16)))
    return out


def nsb_light_rate(flux_x, flux_y, max_fov_radius, max_aperture_radius) -> float:
    """night_sky_background::Light::rate (oracle_nsb_light_rate)."""
    fx, fy = np.ascontiguousarray(flux_x, np.float64), np.ascontiguousarray(flux_y, np.float64)
    return float(_sig_lib().oracle_nsb_light_rate(_ptr(fx, _P_F64), _ptr(fy, _P_F64), len(fx), max_fov_radius,
                                                  max_aperture_radius))


def inject_nsb(offsets, arrival_time, wavelength, truth_id, efficiencies, flux_x, flux_y, rate, exposure_time, seed):
    """(offsets, time, wavelength, truth, start) of the pipelines with the
    night-sky background (oracle_inject_nsb)."""
    off = np.ascontiguousarray(offsets, np.int64)
    ns = len(off) - 1
    t = np.ascontiguousarray(arrival_time, np.float64)
    lam = np.ascontiguousarray(wavelength, np.float64)
    tr = np.ascontiguousarray(truth_id, np.int32)
    eff = np.ascontiguousarray(efficiencies, np.float32)
    fx, fy = np.ascontiguousarray(flux_x, np.float64), np.ascontiguousarray(flux_y, np.float64)
    mean = float(rate) * float(exposure_time) * float(eff.sum(dtype=np.float64)) / max(ns, 1)
    cap = len(t) + int(mean + 10 * np.sqrt(mean + 1) + 64 * (ns + 1))
    oo = np.zeros(ns + 1, np.int64)
    ot, ol, otr = np.empty(cap), np.empty(cap), np.empty(cap, np.int32)
    start = C.c_double()
    k = _sig_lib().oracle_inject_nsb(off.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(t, _P_F64), _ptr(lam, _P_F64),
                                     _ptr(tr, _P_I32), ns, eff.ctypes.data_as(C.POINTER(C.c_float)),
                                     _ptr(fx, _P_F64), _ptr(fy, _P_F64), len(fx), rate, exposure_time, seed,
                                     oo.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(ot, _P_F64), _ptr(ol, _P_F64),
                                     _ptr(otr, _P_I32), cap, C.byref(start))
    if k == -1:
        raise ValueError("bad flux table, sensor rate or wavelength draw")
    if k < 0:
        raise RuntimeError("capacity")
    return oo, ot[:k], ol[:k], otr[:k], start.value


def ref_nsb(offsets, arrival_time, wavelength, truth_id, efficiencies, flux_x, flux_y, exposure_time, seed,
            geometry) -> dict:
    """The reference's inject_nsb_into_photon_pipeline with sensor s drawing
    from Philox stream 9 of (seed, s) (harness `nsb`); geometry = (focal
    length, aperture radius, FoV diameter [deg], pixel FoV [deg], paxels on
    the pixel diagonal) of the light-field sensor whose lixels are the
    sensors."""
    off = np.ascontiguousarray(offsets, np.int64)
    ns = len(off) - 1
    parts = [np.array([ns], np.int32).tobytes(), off.tobytes(),
             np.ascontiguousarray(arrival_time, np.float64).tobytes(),
             np.ascontiguousarray(wavelength, np.float64).tobytes(),
             np.ascontiguousarray(truth_id, np.int32).tobytes(),
             np.ascontiguousarray(efficiencies, np.float32).tobytes(), np.array([len(flux_x)], np.int32).tobytes(),
             np.ascontiguousarray(flux_x, np.float64).tobytes(), np.ascontiguousarray(flux_y, np.float64).tobytes(),
             np.array([exposure_time], np.float64).tobytes(), np.array([seed], np.uint64).tobytes()]
    with tempfile.TemporaryDirectory() as d:
        i, o = Path(d) / "in.bin", Path(d) / "out.bin"
        i.write_bytes(b"".join(parts))
        _run_ref(["nsb", i, o] + [repr(float(v)) if k < 4 else str(int(v)) for k, v in enumerate(geometry)])
        raw = o.read_bytes()
    head = np.frombuffer(raw[:32], np.float64)
    pos = 32
    oo = np.frombuffer(raw[pos:pos + 8 * (ns + 1)], np.int64).copy()
    pos += 8 * (ns + 1)
    k = int(oo[-1])
    t = np.frombuffer(raw[pos:pos + 8 * k], np.float64).copy()
    pos += 8 * k
    lam = np.frombuffer(raw[pos:pos + 8 * k], np.float64).copy()
    pos += 8 * k
    tr = np.frombuffer(raw[pos:pos + 4 * k], np.int32).copy()
    return {"rate": float(head[0]), "start": float(head[1]), "max_fov_radius": float(head[2]),
            "max_aperture_radius": float(head[3]), "offsets": oo, "time": t, "wavelength": lam, "truth": tr}


# ---------------------------------------------------- light-field calibration
CALIBRATION_PHOTON = np.dtype([("reached_sensor", "<u4"), ("lixel_id", "<u4"),
                               ("x_pos_on_principal_aperture", "<f4"), ("y_pos_on_principal_aperture", "<f4"),
                               ("x_tilt_vs_optical_axis", "<f4"), ("y_tilt_vs_optical_axis", "<f4"),
                               ("time_of_flight", "<f4")])
LIXEL_STATISTIC = np.dtype([(k, "<f4") for k in (
    "efficiency", "efficiency_std", "cx_mean", "cx_std", "cy_mean", "cy_std", "x_mean", "x_std", "y_mean", "y_std",
    "time_delay_mean", "time_delay_std")])


def calibration_photons(src: Source, seed: int, photon_id_offset: int, results: dict, num_lixel: int) -> np.ndarray:
    """one_photon's records for oracle results (oracle_calibration_photons)."""
    s = np.ascontiguousarray(results["sensor_id"], np.int32)
    e = np.ascontiguousarray(results["error"], np.uint8)
    t = np.ascontiguousarray(results["time_of_flight"], np.float64)
    out = np.zeros(len(s), CALIBRATION_PHOTON)
    L = lib()
    L.oracle_calibration_photons.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, _P_I32, _P_U8, _P_F64, C.c_int64,
                                             C.c_int32, C.c_void_p]
    L.oracle_calibration_photons(C.byref(src), seed, photon_id_offset, _ptr(s, _P_I32), _ptr(e, _P_U8),
                                 _ptr(t, _P_F64), len(s), num_lixel, out.ctypes.data)
    return out


def lixel_statistics(records: np.ndarray, num_lixel: int, num_blocks: int, photons_per_block: int) -> np.ndarray:
    """LixelStatisticsFiller over the records in order (oracle_lixel_statistics)."""
    r = np.ascontiguousarray(records, CALIBRATION_PHOTON)
    out = np.zeros(num_lixel, LIXEL_STATISTIC)
    L = lib()
    L.oracle_lixel_statistics.restype = C.c_int
    L.oracle_lixel_statistics.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_void_p]
    rc = L.oracle_lixel_statistics(r.ctypes.data, len(r), num_lixel, num_blocks, photons_per_block, out.ctypes.data)
    if rc != 0:
        raise IndexError("a reached lixel id out of range")
    return out


def ref_lixelstats(records: np.ndarray, num_lixel: int, num_blocks: int, photons_per_block: int) -> bytes:
    """The reference's LixelStatisticsFiller over the records in blocks of
    photons_per_block, as calibration::write writes it (harness
    `lixelstats`)."""
    r = np.ascontiguousarray(records, CALIBRATION_PHOTON)
    head = np.array([num_lixel, num_blocks, photons_per_block, len(r)], np.int64).tobytes()
    with tempfile.TemporaryDirectory() as d:
        i, o = Path(d) / "in.bin", Path(d) / "out.bin"
        i.write_bytes(head + r.tobytes())
        _run_ref(["lixelstats", i, o])
        return o.read_bytes()


def fact_photon_stream(offsets, arrival_time, num_pixels, trigger_time=25e-9) -> np.ndarray:
    """FACT's photon stream of one event (oracle_fact_photon_stream)."""
    off = np.ascontiguousarray(offsets, np.int64)
    t = np.ascontiguousarray(arrival_time, np.float64)
    cap = len(t) + num_pixels
    raw = np.zeros(max(cap, 1), np.uint8)
    L = lib()
    L.oracle_fact_photon_stream.restype = C.c_int64
    L.oracle_fact_photon_stream.argtypes = [C.POINTER(C.c_int64), _P_F64, C.c_int32, C.c_double, C.c_void_p,
                                            C.c_int64]
    k = L.oracle_fact_photon_stream(off.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(t, _P_F64), num_pixels,
                                    trigger_time, raw.ctypes.data, cap)
    return raw[:k]


def ref_read_fact_events(path) -> list:
    """The reference's read_SimulationEvent_from_file over a file (harness
    `factphs`): per event (descriptor valid, run, event, reuse, zd, az,
    number_of_photons, raw bytes)."""
    out = _run_ref(["factphs", path])
    events = []
    for line in out.strip().splitlines():
        f = line.split()
        events.append({"valid": int(f[0]), "pass_version": int(f[1]), "event_type": int(f[2]), "run": int(f[3]),
                       "event": int(f[4]), "reuse": int(f[5]), "zd": float(f[6]), "az": float(f[7]),
                       "number_of_photons": int(f[8]), "raw": np.array([int(v) for v in f[9:]], np.uint8)})
    return events


def ref_render(desc, position, rotation, fov, cols, rows, global_illumination, light, sky, frame_colors,
               seed, timing: bool = False):
    """The reference's PinHoleCamera + Tracer per pixel (harness `render`,
    one thread): image rows x cols x 3 float32 (and the seconds of the pixel
    loop with timing=True)."""
    cols_arr = np.ascontiguousarray(frame_colors, np.float32)
    parts = [np.array(position, np.float64).tobytes(), np.array(rotation, np.float64).tobytes(),
             np.array([fov], np.float64).tobytes(), np.array([cols, rows, int(global_illumination), 42],
                                                              np.int32).tobytes(),
             np.array(light, np.float64).tobytes(), np.array(sky, np.float32).tobytes(),
             np.array([len(cols_arr)], np.int32).tobytes(), cols_arr.tobytes(), np.array([seed], np.uint64).tobytes()]
    with tempfile.TemporaryDirectory() as d:
        sp, i, o = Path(d) / "s.bin", Path(d) / "in.bin", Path(d) / "out.bin"
        desc.save(str(sp))
        i.write_bytes(b"".join(parts))
        out = _run_ref(["render", sp, i, o])
        img = np.frombuffer(o.read_bytes(), np.float32).reshape(rows, cols, 3).copy()
    return (img, json.loads(out.strip().splitlines()[-1])["seconds"]) if timing else img


def ref_signal(sensor_id, time_of_flight, wavelength, truth_id, num_sensors, qe_x, qe_y, dark_rate, crosstalk,
               exposure_time, seed, duration, std_dev) -> dict:
    """The reference's signal chain per sensor (harness `signal`)."""
    n = len(sensor_id)
    parts = [np.array([n], np.int64).tobytes(), np.array([num_sensors], np.int32).tobytes(),
             np.ascontiguousarray(sensor_id, np.int32).tobytes(), np.ascontiguousarray(time_of_flight, np.float64).tobytes(),
             np.ascontiguousarray(wavelength, np.float64).tobytes(), np.ascontiguousarray(truth_id, np.int32).tobytes(),
             np.array([len(qe_x)], np.int32).tobytes(), np.ascontiguousarray(qe_x, np.float64).tobytes(),
             np.ascontiguousarray(qe_y, np.float64).tobytes(),
             np.array([dark_rate, crosstalk, exposure_time], np.float64).tobytes(), np.array([seed], np.uint64).tobytes(),
             np.array([duration, std_dev], np.float64).tobytes()]
    with tempfile.TemporaryDirectory() as d:
        i, o = Path(d) / "in.bin", Path(d) / "out.bin"
        i.write_bytes(b"".join(parts))
        _run_ref(["signal", i, o])
        raw = o.read_bytes()
    pos = 0

    def take(dtype, count):
        nonlocal pos
        a = np.frombuffer(raw[pos:pos + np.dtype(dtype).itemsize * count], dtype).copy()
        pos += np.dtype(dtype).itemsize * count
        return a

    out = {}
    out["pipe_offsets"] = take(np.int64, num_sensors + 1)
    m = int(out["pipe_offsets"][-1])
    out["pipe_time"], out["pipe_wavelength"], out["pipe_truth"] = take(np.float64, m), take(np.float64, m), take(np.int32, m)
    out["pulse_offsets"] = take(np.int64, num_sensors + 1)
    k = int(out["pulse_offsets"][-1])
    out["pulse_time"], out["pulse_truth"] = take(np.float64, k), take(np.int32, k)
    out["extracted_offsets"] = take(np.int64, num_sensors + 1)
    x = int(out["extracted_offsets"][-1])
    out["extracted_slice"], out["extracted_truth"] = take(np.int16, x), take(np.int32, x)
    for name in ("photon_stream", "photon_stream_truth"):
        nb = int(take(np.int64, 1)[0])
        out[name] = take(np.uint8, nb)
    return out
