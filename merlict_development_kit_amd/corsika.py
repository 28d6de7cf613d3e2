"""CORSIKA photon bunches to photons on the GPU (mlgpu_corsika_photons).

Replaces merlict::EventIoPhotonFactory (merlict_corsika/PhotonFactory.cpp:8-148)
as the reference's apps apply it to every bunch of an event
(merlict_portal_plenoscope/apps/plenoscope_propagation.cpp:196-203).  A bunch
is 8 float32: x, y [cm], cx, cy, arrival time on the ground [ns], emission
height [cm], weight, wavelength [nm].
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N
from .propagation import PhotonBatch


class BadPhotonWeight(ValueError):
    """A bunch weight outside [0, 16] (EventIoPhotonFactory::BadPhotonWeight)."""


def photons_from_bunches(bunches: torch.Tensor, seed: int = 0, id_offset: int = 0, stream=None):
    """Photons of the bunches (an (n, 8) float32 tensor on the GPU), in bunch
    order, and the bunch index of each (int32).  Bunch b's weight draw is
    Philox stream 5 of (seed, id_offset + b).  Raises BadPhotonWeight like
    the reference."""
    if bunches.dtype != torch.float32 or bunches.dim() != 2 or bunches.shape[1] != 8 \
            or not bunches.is_contiguous() or bunches.device.type != "cuda":
        raise ValueError("bunches must be a contiguous (n, 8) float32 tensor on the GPU")
    n = len(bunches)
    dev = bunches.device
    # every photon count is at most floor(w) + 1
    w = bunches[:, 6].double()
    ok = (w >= 0) & (w <= 16)
    capacity = int((torch.floor(torch.where(ok, w, torch.zeros_like(w))).sum() + n).item()) if n else 0
    arrays = [torch.empty(capacity, dtype=torch.float64, device=dev) for _ in range(7)]
    truth = torch.empty(capacity, dtype=torch.int32, device=dev)
    buf = N.PhotonBuffers(*[a.data_ptr() if capacity else None for a in arrays])
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    m = C.c_int64()
    rc = N.lib().mlgpu_corsika_photons(C.c_void_p(bunches.data_ptr()), n, seed, id_offset, C.byref(buf),
                                       C.c_void_p(truth.data_ptr() if capacity else None), capacity, C.byref(m),
                                       C.c_void_p(st.cuda_stream))
    if rc != N.MLGPU_OK:
        msg = N.lib().mlgpu_last_error().decode(errors="replace")
        if "weight" in msg:
            raise BadPhotonWeight(msg)
        N.check(rc, "mlgpu_corsika_photons")
    k = m.value
    return PhotonBatch(*[a[:k] for a in arrays]), truth[:k]


# ------------------------------------------------------------ EventIO runs
# Fields of the CORSIKA run and event header blocks (merlict_corsika/corsika.cpp:82-125).
def run_number(run_header) -> float:
    return float(run_header[1])


def slope_of_energy_spectrum(run_header) -> float:
    return float(run_header[15])


def energy_range(run_header) -> tuple:
    return float(run_header[16]), float(run_header[17])


def observation_levels(run_header) -> list:
    return [float(run_header[5 + i]) for i in range(int(run_header[4]))]


def event_number(event_header) -> float:
    return float(event_header[1])


def particle_id(event_header) -> float:
    return float(event_header[2])


def total_energy_in_gev(event_header) -> float:
    return float(event_header[3])


def zenith(event_header) -> float:
    return float(event_header[10])


def azimuth(event_header) -> float:
    return float(event_header[11])


def first_interaction_height_in_cm(event_header) -> float:
    return float(event_header[6])


class EventIoEvent:
    """One event of an EventIO run: its 273-float header, its telescope
    offsets (n, 4: toff, xoff, yoff, weight) and its photon bunches (m, 8
    float32, host memory; photons_from_bunches takes them on the GPU)."""

    def __init__(self, header, offsets, bunches):
        self.header = header
        self.telescope_offsets = offsets
        self.bunches = bunches


class EventIoRun:
    """A CORSIKA EventIO run file read event by event (mlgpu_eventio_*,
    eventio::Run in merlict_corsika/eventio.h:124-146): ``header`` (273
    floats), ``input_card`` (bytes), ``telescope_positions`` (n, 4: x, y, z,
    r); ``has_events()`` / ``next_event()`` or iteration; ``end`` (the run
    end block) once the last event has been read."""

    def __init__(self, path):
        import numpy as np
        self._np = np
        self._h = C.c_void_p()
        N.check(N.lib().mlgpu_eventio_open(str(path).encode(), C.byref(self._h)), "mlgpu_eventio_open")
        self.header = np.zeros(273, np.float32)
        ntel, ncard = C.c_int32(), C.c_int64()
        N.check(N.lib().mlgpu_eventio_run_header(self._h, self.header.ctypes.data, C.byref(ntel), C.byref(ncard)),
                "mlgpu_eventio_run_header")
        self.telescope_positions = np.zeros((ntel.value, 4), np.float32)
        card = C.create_string_buffer(max(ncard.value, 1))
        N.check(N.lib().mlgpu_eventio_run_info(self._h, self.telescope_positions.ctypes.data, card),
                "mlgpu_eventio_run_info")
        self.input_card = card.raw[:ncard.value]

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            N.lib().mlgpu_eventio_close(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def has_events(self) -> bool:
        return bool(N.lib().mlgpu_eventio_has_events(self._h))

    def next_event(self) -> EventIoEvent:
        np = self._np
        header = np.zeros(273, np.float32)
        no, nb = C.c_int32(), C.c_int64()
        N.check(N.lib().mlgpu_eventio_next_event(self._h, header.ctypes.data, C.byref(no), C.byref(nb)),
                "mlgpu_eventio_next_event")
        offsets = np.zeros((no.value, 4), np.float32)
        bunches = np.zeros((nb.value, 8), np.float32)
        N.check(N.lib().mlgpu_eventio_event_data(self._h, offsets.ctypes.data, bunches.ctypes.data),
                "mlgpu_eventio_event_data")
        return EventIoEvent(header, offsets, bunches)

    def __iter__(self):
        while self.has_events():
            yield self.next_event()

    @property
    def end(self):
        out = self._np.zeros(273, self._np.float32)
        N.check(N.lib().mlgpu_eventio_run_end(self._h, out.ctypes.data), "mlgpu_eventio_run_end")
        return out


class EventTapeRun:
    """A CORSIKA event tape (a tar of RUNH, EVTH and Cherenkov bunch blocks)
    read event by event (mlgpu_event_tape_*, event_tape::Run in
    merlict_corsika/event_tape.h:23-36): ``header`` (273 floats),
    ``has_events()`` / ``next_event()`` (header and (n, 8) float32 bunches)
    or iteration."""

    def __init__(self, path):
        import numpy as np
        self._np = np
        self._h = C.c_void_p()
        N.check(N.lib().mlgpu_event_tape_open(str(path).encode(), C.byref(self._h)), "mlgpu_event_tape_open")
        self.header = np.zeros(273, np.float32)
        N.check(N.lib().mlgpu_event_tape_run_header(self._h, self.header.ctypes.data), "mlgpu_event_tape_run_header")

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            N.lib().mlgpu_event_tape_close(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def has_events(self) -> bool:
        return bool(N.lib().mlgpu_event_tape_has_events(self._h))

    def next_event(self) -> EventIoEvent:
        np = self._np
        header = np.zeros(273, np.float32)
        nb = C.c_int64()
        N.check(N.lib().mlgpu_event_tape_next_event(self._h, header.ctypes.data, C.byref(nb)),
                "mlgpu_event_tape_next_event")
        bunches = np.zeros((nb.value, 8), np.float32)
        if nb.value:
            N.check(N.lib().mlgpu_event_tape_event_bunches(self._h, bunches.ctypes.data),
                    "mlgpu_event_tape_event_bunches")
        return EventIoEvent(header, np.zeros((0, 4), np.float32), bunches)

    def __iter__(self):
        while self.has_events():
            yield self.next_event()
