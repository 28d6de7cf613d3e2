"""ctypes binding of ``libmlgpu.so`` (the C ABI declared in ``include/mlgpu.h``).

The library is built in-tree by :mod:`merlict_development_kit_amd.build`.
Loading fails loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmlgpu.so"
# the MLGPU_ABI_VERSION of include/mlgpu.h these bindings follow
ABI_VERSION = 2

MLGPU_OK = 0
MLGPU_NUM_PARAMS = 6
MLGPU_NUM_SURFACE_SLOTS = 6

KIND = {
    "Frame": 0,
    "SphereCapWithHexagonalBound": 1,
    "SphereCapWithCylinderBound": 2,
    "SphereCapWithRectangularBound": 3,
    "Plane": 4,
    "Disc": 5,
    "HexPlane": 6,
    "Annulus": 7,
    "PlaneDualSphericalBound": 8,
    "Cylinder": 9,
    "Triangle": 10,
    "Sphere": 11,
    "EllipticalCapWithHexagonalBound": 12,
}

# merlict::Interaction (merlict/RayForPropagation.h:14-28)
PRODUCTION = 0
ABSORPTION_IN_VOID = 1
ABSORPTION_IN_MEDIUM = 2
ABSORPTION_ON_SURFACE = 3
FRESNEL_REFLECTION_ON_SURFACE = 4
REFLECTION_ON_SURFACE = 5
REFRACTION_TO_OUTSIDE = 6
REFRACTION_TO_INSIDE = 7

SCHEDULE_AUTO = 0
SCHEDULE_FUSED = 1
SCHEDULE_WAVEFRONT = 2
SCHEDULE_PERSISTENT = 3

PHOTON_OK = 0
PHOTON_BAD_WAVELENGTH = 1
PHOTON_OUT_OF_TABLE = 2

_P_I32 = C.POINTER(C.c_int32)
_P_F64 = C.POINTER(C.c_double)


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_frames", C.c_int32),
        ("mother", _P_I32),
        ("pos", _P_F64),
        ("rot_mode", _P_I32),
        ("rot", _P_F64),
        ("kind", _P_I32),
        ("params", _P_F64),
        ("surface", _P_I32),
        ("allowed_frame", _P_I32),
        ("sensor_id", _P_I32),
        ("n_functions", C.c_int32),
        ("func_start", _P_I32),
        ("func_x", _P_F64),
        ("func_y", _P_F64),
    ]


class Photons(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "support_x", "support_y", "support_z", "direction_x", "direction_y", "direction_z", "wavelength")]


class PhotonBuffers(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "support_x", "support_y", "support_z", "direction_x", "direction_y", "direction_z", "wavelength")]


class Source(C.Structure):
    _fields_ = [
        ("support", C.c_int32),
        ("direction", C.c_int32),
        ("radius", C.c_double),
        ("angle", C.c_double),
        ("lambda_min", C.c_double),
        ("lambda_max", C.c_double),
        ("start_distance", C.c_double),
        ("rot", C.c_double * 9),
        ("pos", C.c_double * 3),
        ("angle_width", C.c_double),
        ("spectrum", C.c_int32),
        ("pad", C.c_int32),
    ]


class Bunch(C.Structure):
    _fields_ = [
        ("source", Source),
        ("seed", C.c_uint64),
        ("num_photons", C.c_int64),
    ]


class SurrogateDomain(C.Structure):
    _fields_ = [(name, C.c_double) for name in (
        "support_radius", "max_angle", "lambda_min", "lambda_max", "image_scale", "position_scale", "time_offset",
        "time_scale")]


class HexCells(C.Structure):
    _fields_ = [
        ("a1", C.c_double * 2),
        ("a2", C.c_double * 2),
        ("origin", C.c_double * 2),
        ("radius", C.c_int32),
        ("cell", C.c_void_p),
        ("centre", C.c_void_p),
        ("inner_radius", C.c_double),
    ]


class HybridConfig(C.Structure):
    _fields_ = [("max_entropy", C.c_double), ("seed", C.c_uint64)]


class Histogram(C.Structure):
    _fields_ = [
        ("counts", C.c_void_p),
        ("num_sensors", C.c_int32),
        ("num_bins", C.c_int32),
        ("t_start", C.c_double),
        ("bin_width", C.c_double),
        ("jitter_sigma", C.c_double),
        ("seed", C.c_uint64),
        ("photon_id_offset", C.c_uint64),
    ]


class Results(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in (
        "final_frame", "sensor_id", "final_interaction", "num_interactions", "error", "time_of_flight",
        "x", "y", "cx", "cy", "accumulated_distance")]


class Pipelines(C.Structure):
    _fields_ = [("offsets", C.c_void_p), ("arrival_time", C.c_void_p), ("wavelength", C.c_void_p),
                ("truth_id", C.c_void_p), ("capacity", C.c_int64)]


class Pulses(C.Structure):
    _fields_ = [("offsets", C.c_void_p), ("arrival_time", C.c_void_p), ("truth_id", C.c_void_p),
                ("capacity", C.c_int64)]


class Converter(C.Structure):
    _fields_ = [("qe_x", C.c_void_p), ("qe_y", C.c_void_p), ("qe_n", C.c_int32), ("dark_rate", C.c_double),
                ("crosstalk", C.c_double), ("exposure_time", C.c_double), ("seed", C.c_uint64)]


class Nsb(C.Structure):
    _fields_ = [("flux_x", C.c_void_p), ("flux_y", C.c_void_p), ("flux_n", C.c_int32), ("rate", C.c_double),
                ("exposure_time", C.c_double), ("seed", C.c_uint64)]


class PinholeCamera(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("rotation", C.c_double * 3), ("field_of_view", C.c_double),
                ("num_cols", C.c_int32), ("num_rows", C.c_int32)]


class RenderConfig(C.Structure):
    _fields_ = [("global_illumination", C.c_int32), ("max_interaction_depth", C.c_int32),
                ("incoming_direction", C.c_double * 3), ("sky_color", C.c_float * 3), ("pad", C.c_float)]


class Config(C.Structure):
    _fields_ = [
        ("max_num_interactions_per_photon", C.c_uint32),
        ("seed", C.c_uint64),
        ("photon_id_offset", C.c_uint64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("schedule", C.c_int32),
        ("trace_budget", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("tail_level", C.c_int32),
        ("shade_lanes", C.c_int32),
    ]


class History(C.Structure):
    _fields_ = [("bounces", C.c_void_p), ("capacity", C.c_int32), ("num_bounces", C.c_void_p),
                ("final_ray", C.c_void_p)]


# numpy layout of mlgpu_bounce (64 bytes)
BOUNCE_FIELDS = [("position", "<f8", (3,)), ("normal", "<f8", (3,)), ("distance", "<f8"), ("frame", "<i4"),
                 ("interaction", "u1"), ("from_outside", "u1"), ("reserved", "u1", (2,))]


class Hits(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in ("frame", "distance", "position_x", "position_y", "position_z")]


# every function include/mlgpu.h declares, with its ctypes signature
SIGNATURES = {
    "mlgpu_last_error": (C.c_char_p, []),
    "mlgpu_abi_version": (C.c_int, []),
    "mlgpu_desc_destroy": (None, [C.c_void_p]),
    "mlgpu_desc_view": (C.c_int, [C.c_void_p, C.POINTER(SceneDesc)]),
    "mlgpu_desc_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "mlgpu_desc_frame_poses": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_double)]),
    "mlgpu_desc_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "mlgpu_desc_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "mlgpu_build_lens_scene": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    "mlgpu_build_fact_scene": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "mlgpu_build_plenoscope_scene": (C.c_int, [C.c_double] * 7 + [C.c_int, C.c_double, C.c_double, C.c_double,
                                                                   C.POINTER(C.c_void_p)]),
    "mlgpu_scene_create": (C.c_int, [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    "mlgpu_scene_destroy": (None, [C.c_void_p]),
    "mlgpu_scene_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "mlgpu_scene_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    "mlgpu_scene_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Photons), C.c_int64, C.POINTER(Config),
                                       C.c_int32, C.POINTER(Tuning), C.POINTER(C.c_double), C.c_void_p]),
    "mlgpu_scene_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mlgpu_propagate": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Photons), C.c_int64, C.POINTER(Config),
                                  C.POINTER(Results), C.c_void_p]),
    "mlgpu_propagate_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Photons), C.c_int64,
                                          C.POINTER(Config), C.POINTER(Results), C.c_void_p]),
    "mlgpu_propagate_history": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Photons), C.c_int64,
                                          C.POINTER(Config), C.POINTER(Results), C.POINTER(History), C.c_void_p]),
    "mlgpu_generate_photons": (C.c_int, [C.POINTER(Source), C.c_int64, C.c_uint64, C.c_uint64,
                                         C.POINTER(PhotonBuffers), C.c_void_p]),
    "mlgpu_propagate_bunches": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Bunch), C.c_int32, C.POINTER(Config),
                                          C.POINTER(Histogram), C.POINTER(Results), C.c_void_p]),
    "mlgpu_surrogate_num_params": (C.c_int64, []),
    "mlgpu_surrogate_create": (C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    "mlgpu_surrogate_destroy": (None, [C.c_void_p]),
    "mlgpu_surrogate_get_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mlgpu_surrogate_set_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mlgpu_event_tape_open": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "mlgpu_event_tape_close": (None, [C.c_void_p]),
    "mlgpu_event_tape_run_header": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mlgpu_event_tape_has_events": (C.c_int, [C.c_void_p]),
    "mlgpu_event_tape_next_event": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]),
    "mlgpu_event_tape_event_bunches": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mlgpu_eventio_open": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "mlgpu_eventio_close": (None, [C.c_void_p]),
    "mlgpu_eventio_run_header": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "mlgpu_eventio_run_info": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mlgpu_eventio_has_events": (C.c_int, [C.c_void_p]),
    "mlgpu_eventio_next_event": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "mlgpu_eventio_event_data": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mlgpu_eventio_run_end": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mlgpu_surrogate_set_inference": (C.c_int, [C.c_void_p, C.c_int32]),
    "mlgpu_surrogate_get_inference": (C.c_int32, [C.c_void_p]),
    "mlgpu_surrogate_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "mlgpu_surrogate_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
    "mlgpu_surrogate_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_void_p,
                                             C.c_void_p]),
    "mlgpu_surrogate_set_region": (C.c_int, [C.c_void_p, C.POINTER(SurrogateDomain), C.POINTER(HexCells)]),
    "mlgpu_surrogate_set_survival_scale": (C.c_int, [C.c_void_p, C.c_float]),
    "mlgpu_surrogate_prepare": (C.c_int, [C.POINTER(SurrogateDomain), C.POINTER(Photons), C.POINTER(Results),
                                          C.c_void_p, C.POINTER(Results), C.c_void_p, C.c_int64, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    "mlgpu_propagate_hybrid": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SurrogateDomain), C.POINTER(HexCells),
                                         C.c_int32, C.POINTER(Photons), C.c_int64, C.POINTER(Config),
                                         C.POINTER(HybridConfig), C.POINTER(Histogram), C.POINTER(C.c_int64),
                                         C.c_void_p]),
    "mlgpu_bunch_plan_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, C.POINTER(Config),
                                          C.POINTER(Histogram), C.POINTER(C.c_void_p)]),
    "mlgpu_bunch_plan_run": (C.c_int, [C.c_void_p, C.POINTER(Bunch), C.c_int32, C.POINTER(Config), C.c_void_p]),
    "mlgpu_bunch_plan_results": (C.c_int, [C.c_void_p, C.POINTER(Results)]),
    "mlgpu_bunch_plan_destroy": (None, [C.c_void_p]),
    "mlgpu_generate_bunches": (C.c_int, [C.POINTER(Bunch), C.c_int32, C.POINTER(PhotonBuffers), C.c_int64,
                                         C.c_void_p]),
    "mlgpu_sensor_pipelines": (C.c_int, [C.POINTER(Results), C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                         C.POINTER(Pipelines), C.POINTER(C.c_int64), C.c_void_p]),
    "mlgpu_photoelectric_convert": (C.c_int, [C.POINTER(Pipelines), C.c_int32, C.POINTER(Converter),
                                              C.POINTER(Pulses), C.POINTER(C.c_int64), C.c_void_p]),
    "mlgpu_extract_pulses": (C.c_int, [C.POINTER(Pulses), C.c_int32, C.c_double, C.c_double, C.c_uint64, C.c_void_p,
                                       C.c_void_p]),
    "mlgpu_photon_stream": (C.c_int, [C.POINTER(Pulses), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                      C.POINTER(C.c_int64), C.c_void_p]),
    "mlgpu_fact_photon_stream": (C.c_int, [C.POINTER(Pulses), C.c_int32, C.c_double, C.c_void_p, C.c_int64,
                                           C.POINTER(C.c_int64), C.c_void_p]),
    "mlgpu_inject_nsb": (C.c_int, [C.POINTER(Pipelines), C.c_int32, C.c_void_p, C.POINTER(Nsb), C.POINTER(Pipelines),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_void_p]),
    "mlgpu_nsb_light_rate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_double,
                                       C.POINTER(C.c_double)]),
    "mlgpu_render_pinhole": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(PinholeCamera), C.POINTER(RenderConfig),
                                       C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p, C.POINTER(C.c_int64),
                                       C.c_void_p]),
    "mlgpu_calibration_photons": (C.c_int, [C.POINTER(Source), C.c_uint64, C.c_uint64, C.POINTER(Results), C.c_int64,
                                            C.c_int32, C.c_void_p, C.c_void_p]),
    "mlgpu_lixel_filler_create": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)]),
    "mlgpu_lixel_filler_fill_block": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "mlgpu_lixel_filler_statistics": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "mlgpu_lixel_filler_destroy": (None, [C.c_void_p]),
    "mlgpu_corsika_photons": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint64, C.c_uint64, C.POINTER(PhotonBuffers),
                                        C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_void_p]),
    "mlgpu_histogram_arrivals": (C.c_int, [C.POINTER(Results), C.c_int64, C.POINTER(Histogram), C.c_void_p]),
    "mlgpu_count_errors": (C.c_int, [C.POINTER(Results), C.c_int64, C.c_void_p, C.POINTER(C.c_int64)]),
    "mlgpu_first_hit": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Photons), C.c_int64, C.POINTER(Hits),
                                  C.c_void_p]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded native library (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing; build it with `python -m merlict_development_kit_amd.build`")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        if handle.mlgpu_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {handle.mlgpu_abi_version()}, these bindings "
                               f"expect {ABI_VERSION}; rebuild it")
        _lib = handle
    return _lib


class MlgpuError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = lib().mlgpu_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with status {status}: {msg}")
        self.status = status


def check(status: int, where: str) -> None:
    if status != MLGPU_OK:
        raise MlgpuError(status, where)
