"""The sensor signal chain on the GPU (merlict_signal_processing):
per-sensor photon pipelines sorted by arrival time, the photo-electric
converter, the night-sky-background injection and the single-photon pulse
extraction (include/mlgpu.h)."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np
import torch

from . import _native as N


def _st(stream, dev):
    s = stream if stream is not None else torch.cuda.current_stream(dev)
    return C.c_void_p(s.cuda_stream)


@dataclass
class Pipelines:
    """Pipeline s: entries offsets[s] .. offsets[s + 1] - 1."""
    offsets: torch.Tensor        # int64, num_sensors + 1
    arrival_time: torch.Tensor   # float64
    wavelength: torch.Tensor     # float64
    truth_id: torch.Tensor       # int32

    def c_struct(self) -> N.Pipelines:
        return N.Pipelines(self.offsets.data_ptr(), self.arrival_time.data_ptr(), self.wavelength.data_ptr(),
                           self.truth_id.data_ptr(), len(self.arrival_time))


@dataclass
class Pulses:
    offsets: torch.Tensor        # int64, num_sensors + 1
    arrival_time: torch.Tensor   # float64
    truth_id: torch.Tensor       # int32

    def c_struct(self) -> N.Pulses:
        return N.Pulses(self.offsets.data_ptr(), self.arrival_time.data_ptr(), self.truth_id.data_ptr(),
                        len(self.arrival_time))


def sensor_pipelines(sensor_id: torch.Tensor, error: torch.Tensor, time_of_flight: torch.Tensor,
                     wavelength: torch.Tensor, num_sensors: int, truth_id: Optional[torch.Tensor] = None,
                     stream=None) -> Pipelines:
    """Per-sensor pipelines of the detected photons (mlgpu_sensor_pipelines);
    the arrays are per photon, on the GPU (a PropagationResult's sensor_id,
    error and time_of_flight with the photons' wavelengths)."""
    dev = sensor_id.device
    n = len(sensor_id)
    r = N.Results()
    r.sensor_id, r.error, r.time_of_flight = sensor_id.data_ptr(), error.data_ptr(), time_of_flight.data_ptr()
    out = Pipelines(torch.zeros(num_sensors + 1, dtype=torch.int64, device=dev),
                    torch.empty(n, dtype=torch.float64, device=dev), torch.empty(n, dtype=torch.float64, device=dev),
                    torch.empty(n, dtype=torch.int32, device=dev))
    m = C.c_int64()
    N.check(N.lib().mlgpu_sensor_pipelines(C.byref(r), C.c_void_p(wavelength.data_ptr()),
                                           C.c_void_p(truth_id.data_ptr() if truth_id is not None else None), n,
                                           num_sensors, C.byref(out.c_struct()), C.byref(m), _st(stream, dev)),
            "mlgpu_sensor_pipelines")
    k = m.value
    return Pipelines(out.offsets, out.arrival_time[:k], out.wavelength[:k], out.truth_id[:k])


def photoelectric_convert(pipes: Pipelines, num_sensors: int, qe_x, qe_y, dark_rate: float = 0.0,
                          crosstalk: float = 0.0, exposure_time: float = 0.0, seed: int = 0,
                          stream=None) -> Pulses:
    """Pulses of every pipeline (mlgpu_photoelectric_convert)."""
    dev = pipes.offsets.device
    qx = np.ascontiguousarray(qe_x, np.float64)
    qy = np.ascontiguousarray(qe_y, np.float64)
    conv = N.Converter(qx.ctypes.data, qy.ctypes.data, len(qx), dark_rate, crosstalk, exposure_time, seed)
    m = len(pipes.arrival_time)
    cap = int(m / max(1e-9, 1.0 - crosstalk)) * 2 + int(dark_rate * exposure_time * num_sensors * 2 / max(
        1e-9, 1.0 - crosstalk)) + 64 * (num_sensors + 1)
    for _ in range(2):
        out = Pulses(torch.zeros(num_sensors + 1, dtype=torch.int64, device=dev),
                     torch.empty(cap, dtype=torch.float64, device=dev), torch.empty(cap, dtype=torch.int32, device=dev))
        k = C.c_int64()
        rc = N.lib().mlgpu_photoelectric_convert(C.byref(pipes.c_struct()), num_sensors, C.byref(conv),
                                                 C.byref(out.c_struct()), C.byref(k), _st(stream, dev))
        if rc == N.MLGPU_OK:
            return Pulses(out.offsets, out.arrival_time[:k.value], out.truth_id[:k.value])
        if k.value > cap:  # the count is known: once more with room for it
            cap = k.value
            continue
        N.check(rc, "mlgpu_photoelectric_convert")
    N.check(rc, "mlgpu_photoelectric_convert")


def extract_pulses(pulses: Pulses, num_sensors: int, time_slice_duration: float, arrival_time_std: float,
                   seed: int = 0, stream=None) -> torch.Tensor:
    """Time slice of every pulse, -1 when outside the 100 slices
    (mlgpu_extract_pulses)."""
    dev = pulses.offsets.device
    out = torch.empty(len(pulses.arrival_time), dtype=torch.int16, device=dev)
    N.check(N.lib().mlgpu_extract_pulses(C.byref(pulses.c_struct()), num_sensors, time_slice_duration,
                                         arrival_time_std, seed, C.c_void_p(out.data_ptr()), _st(stream, dev)),
            "mlgpu_extract_pulses")
    return out


def nsb_light_rate(flux_x, flux_y, max_fov_radius: float, max_aperture_radius: float) -> float:
    """Rate [1/s] of the night-sky light over the whole light-field sensor
    (night_sky_background::Light, Light.cpp:9-28; mlgpu_nsb_light_rate):
    max_fov_radius is half the sensor's max FoV diameter [rad]."""
    fx = np.ascontiguousarray(flux_x, np.float64)
    fy = np.ascontiguousarray(flux_y, np.float64)
    r = C.c_double()
    N.check(N.lib().mlgpu_nsb_light_rate(fx.ctypes.data, fy.ctypes.data, len(fx), max_fov_radius,
                                         max_aperture_radius, C.byref(r)), "mlgpu_nsb_light_rate")
    return r.value


def inject_nsb(pipes: Pipelines, num_sensors: int, efficiencies: torch.Tensor, flux_x, flux_y, rate: float,
               exposure_time: float, seed: int = 0, stream=None):
    """(pipelines with the night-sky background, exposure start time)
    (inject_nsb_into_photon_pipeline, Injector.cpp:11-107;
    mlgpu_inject_nsb).  efficiencies: float32 per sensor on the GPU."""
    dev = pipes.offsets.device
    if efficiencies.dtype != torch.float32 or efficiencies.device != dev or len(efficiencies) != num_sensors:
        raise ValueError("efficiencies must be num_sensors float32 values on the pipelines' device")
    eff = efficiencies.contiguous()
    fx = np.ascontiguousarray(flux_x, np.float64)
    fy = np.ascontiguousarray(flux_y, np.float64)
    nsb = N.Nsb(fx.ctypes.data, fy.ctypes.data, len(fx), rate, exposure_time, seed)
    m = len(pipes.arrival_time)
    mean = rate * exposure_time * float(eff.double().sum().item()) / max(num_sensors, 1)
    # a first guess (a larger count comes back and is allocated exactly)
    cap = m + int(min(mean + 10.0 * math.sqrt(mean + 1.0), float(1 << 26))) + 16 * (num_sensors + 1)
    for _ in range(2):
        out = Pipelines(torch.zeros(num_sensors + 1, dtype=torch.int64, device=dev),
                        torch.empty(cap, dtype=torch.float64, device=dev),
                        torch.empty(cap, dtype=torch.float64, device=dev),
                        torch.empty(cap, dtype=torch.int32, device=dev))
        k, start = C.c_int64(), C.c_double()
        rc = N.lib().mlgpu_inject_nsb(C.byref(pipes.c_struct()), num_sensors, C.c_void_p(eff.data_ptr()),
                                      C.byref(nsb), C.byref(out.c_struct()), C.byref(k), C.byref(start),
                                      _st(stream, dev))
        if rc == N.MLGPU_OK:
            n = k.value
            return Pipelines(out.offsets, out.arrival_time[:n], out.wavelength[:n], out.truth_id[:n]), start.value
        if k.value > cap:  # the count is known: once more with room for it
            cap = k.value
            continue
        N.check(rc, "mlgpu_inject_nsb")
    N.check(rc, "mlgpu_inject_nsb")


NEXT_CHANNEL_MARKER = 255   # merlict_signal_processing/PhotonStream.h:15
NUMBER_TIME_SLICES = 100    # pulse_extraction.h:13


def photon_stream(pulses: Pulses, slices: torch.Tensor, num_sensors: int, stream=None):
    """(symbols uint8, truth int32) on the GPU: the kept pulses' slices
    channel by channel with a marker between channels, and their truth ids
    (mlgpu_photon_stream)."""
    dev = pulses.offsets.device
    if slices.dtype != torch.int16 or len(slices) != len(pulses.arrival_time):
        raise ValueError("slices must be one int16 per pulse")
    cap = len(slices) + num_sensors
    sym = torch.empty(max(cap, 1), dtype=torch.uint8, device=dev)
    truth = torch.empty(max(len(slices), 1), dtype=torch.int32, device=dev)
    k = C.c_int64()
    N.check(N.lib().mlgpu_photon_stream(C.byref(pulses.c_struct()), C.c_void_p(slices.data_ptr() if len(slices) else None),
                                        num_sensors, C.c_void_p(sym.data_ptr()), C.c_void_p(truth.data_ptr()), cap,
                                        C.byref(k), _st(stream, dev)), "mlgpu_photon_stream")
    n = k.value
    kept = n - num_sensors + 1 if n > 0 else 0
    return sym[:n], truth[:kept]


def write_photon_stream(path, symbols, num_channels: int, time_slice_duration: float) -> None:
    """PhotonStream::write (PhotonStream.cpp:15-58): float32 slice duration,
    uint32 channels, uint32 100 slices, uint32 symbols, then the symbols."""
    sym = symbols.cpu().numpy() if isinstance(symbols, torch.Tensor) else np.asarray(symbols, np.uint8)
    head = np.array([time_slice_duration], "<f4").tobytes() + np.array(
        [num_channels, NUMBER_TIME_SLICES, len(sym)], "<u4").tobytes()
    Path(path).write_bytes(head + np.ascontiguousarray(sym, np.uint8).tobytes())


def write_simulation_truth(path, truth) -> None:
    """PhotonStream::write_simulation_truth (PhotonStream.cpp:60-78): int32 per pulse."""
    t = truth.cpu().numpy() if isinstance(truth, torch.Tensor) else np.asarray(truth)
    Path(path).write_bytes(np.ascontiguousarray(t, "<i4").tobytes())


def read_photon_stream(path):
    """PhotonStream::read (PhotonStream.cpp:80-130): (time slice duration,
    list of per-channel uint8 slice arrays)."""
    raw = Path(path).read_bytes()
    dur = float(np.frombuffer(raw[:4], "<f4")[0])
    _, slices, n = (int(v) for v in np.frombuffer(raw[4:16], "<u4"))
    if slices != NUMBER_TIME_SLICES:
        raise ValueError(f"expected {NUMBER_TIME_SLICES} time slices, got {slices}")
    sym = np.frombuffer(raw[16:16 + n], np.uint8)
    cut = np.nonzero(sym == NEXT_CHANNEL_MARKER)[0]
    channels = np.split(sym, cut)
    channels = [channels[0]] + [c[1:] for c in channels[1:]] if n > 0 else []
    return dur, channels


# merlict_fact/apps/photon_stream.h:29-37
FACT_SLICE_OFFSET = 30
FACT_PASS_VERSION = 4
FACT_SIMULATION_EVENT = 1


def fact_photon_stream(pulses: Pulses, num_pixels: int = 1440, trigger_time: float = 25e-9,
                       stream=None) -> torch.Tensor:
    """The raw bytes of FACT's photon stream of one event (uint8 on the GPU):
    the pulses' slices relative to the event's median arrival time, pixel
    by pixel, each pixel ended by 255 (mlgpu_fact_photon_stream)."""
    dev = pulses.offsets.device
    cap = len(pulses.arrival_time) + num_pixels
    raw = torch.empty(max(cap, 1), dtype=torch.uint8, device=dev)
    k = C.c_int64()
    N.check(N.lib().mlgpu_fact_photon_stream(C.byref(pulses.c_struct()), num_pixels, trigger_time,
                                             C.c_void_p(raw.data_ptr()), cap, C.byref(k), _st(stream, dev)),
            "mlgpu_fact_photon_stream")
    return raw[:k.value]


def write_fact_simulation_event(fout, raw, run: int, event: int, reuse: int = 0, zd: float = 0.0,
                                az: float = 0.0) -> None:
    """append_SimulationEvent_to_file (merlict_fact/apps/photon_stream.h:398-406):
    descriptor 'phs', pass version, event type; run, event, reuse; zenith
    distance and azimuth; the stream's length, its bytes and no saturated
    pixels."""
    r = raw.cpu().numpy() if isinstance(raw, torch.Tensor) else np.asarray(raw, np.uint8)
    fout.write(b"phs" + bytes([FACT_PASS_VERSION, FACT_SIMULATION_EVENT]))
    fout.write(np.array([run, event, reuse], "<u4").tobytes())
    fout.write(np.array([zd, az], "<f4").tobytes())
    fout.write(np.array([len(r)], "<u4").tobytes())
    fout.write(np.ascontiguousarray(r, np.uint8).tobytes())
    fout.write(np.array([0], "<u2").tobytes())


def read_fact_simulation_events(path) -> list:
    """read_SimulationEvent_from_file (merlict_fact/apps/photon_stream.h:386-396)
    over a whole file: one dict per event."""
    raw = Path(path).read_bytes()
    events, pos = [], 0
    while pos < len(raw):
        if raw[pos:pos + 3] != b"phs" or raw[pos + 4] != FACT_SIMULATION_EVENT:
            raise ValueError("not a FACT simulation event")
        run, event, reuse = (int(v) for v in np.frombuffer(raw[pos + 5:pos + 17], "<u4"))
        zd, az = (float(v) for v in np.frombuffer(raw[pos + 17:pos + 25], "<f4"))
        n = int(np.frombuffer(raw[pos + 25:pos + 29], "<u4")[0])
        stream = np.frombuffer(raw[pos + 29:pos + 29 + n], np.uint8).copy()
        sat = int(np.frombuffer(raw[pos + 29 + n:pos + 31 + n], "<u2")[0])
        events.append({"pass_version": raw[pos + 3], "run": run, "event": event, "reuse": reuse, "zd": zd,
                       "az": az, "raw": stream})
        # append_PhotonStream_to_file writes `sat` bytes of the saturated pixels'
        # list (photon_stream.h:336-340)
        pos += 31 + n + sat
    return events
