"""CORSIKA EventIO runs on the host (mlgpu_eventio_*): the reference's own
tests of eventio::Run (merlict_corsika/tests/EventIoTest.cpp) on their file
telescope.dat (copied to tests/golden/corsika_telescope.dat), the reader's
error paths, and every bunch bit for bit against the reference's reader
(the reference harness's `corsika-extract`, eventio::Run) where the harness
is built.  The reader is host code: these tests need no GPU."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import pyoracle as O

GOLD = Path(__file__).parent / "golden" / "corsika_telescope.dat"


def _run(path=GOLD):
    from merlict_development_kit_amd.corsika import EventIoRun
    return EventIoRun(path)


def test_eventio_run_header():
    """EventIoTest.cpp:36-80: the run header block starts with 'RUNH', one
    telescope at (0, 0, 500) of radius 500, the input card, run 7 with an
    E^-2.7 spectrum from 1 to 50 TeV and one observation level at 2.2 km."""
    from merlict_development_kit_amd import corsika as K
    r = _run()
    assert r.header[0] == np.frombuffer(b"RUNH", np.float32)[0]
    np.testing.assert_array_equal(r.telescope_positions, [[0.0, 0.0, 500.0, 500.0]])
    assert r.input_card[100:101] == b" " and len(r.input_card) > 100
    assert abs(K.run_number(r.header) - 7.0) <= 1e-6
    assert abs(K.slope_of_energy_spectrum(r.header) + 2.7) <= 1e-6
    lo, hi = K.energy_range(r.header)
    assert abs(lo - 1000.0) <= 1e-6 and abs(hi - 50000.0) <= 1e-6
    levels = K.observation_levels(r.header)
    assert len(levels) == 1 and abs(levels[0] - 220000.0) <= 1e-6


def test_eventio_first_event():
    """EventIoTest.cpp:86-186: the first event's telescope offset, event
    number, primary and energy; 42,629 bunches; bunches 2-4 (x, y, cx, cy,
    time, weight and wavelength to Catch's default relative margin, the
    emission height within 10 cm, as the reference's test)."""
    from merlict_development_kit_amd import corsika as K
    r = _run()
    assert r.has_events()
    ev = r.next_event()
    assert ev.telescope_offsets.shape == (1, 4)
    toff, xoff, yoff, weight = ev.telescope_offsets[0]
    assert abs(toff - 379489.3125) <= 1e-6 and abs(xoff) <= 1e-6 and abs(yoff + 6589.96044922) <= 1e-6
    assert weight == 1.0
    assert K.event_number(ev.header) == 1.0 and K.particle_id(ev.header) == 1.0
    assert abs(K.total_energy_in_gev(ev.header) - 2745.3125) <= 1e-6
    assert ev.bunches.shape == (42629, 8) and ev.bunches.dtype == np.float32
    want = np.array([
        [-1.899999976158142, 286.20001220703125, 0.23173333704471588, 0.017733333632349968, -13.0, 1458816.0, 1.0,
         -351.0],
        [43.5, 483.0, 0.23313333094120026, 0.0351666659116745, -10.199999809265137, 1276440.0, 0.9899999499320984,
         -373.0],
        [7.300000190734863, -457.1000061035156, 0.23346666991710663, 0.03263333439826965, -11.699999809265137,
         1358314.625, 0.9899999499320984, -523.0]])
    got = ev.bunches[2:5].astype(np.float64)
    eps = 100 * np.finfo(np.float32).eps
    cols = [0, 1, 2, 3, 4, 6, 7]
    assert np.all(np.abs(got[:, cols] - want[:, cols]) <= eps * (1 + np.abs(want[:, cols])))
    assert np.all(np.abs(got[:, 5] - want[:, 5]) <= 10.0)


def test_eventio_whole_run():
    """EventIoTest.cpp:188-206: every event of the run reads (the file holds
    one); after it the run end is there and no event is left."""
    r = _run()
    events = list(r)
    assert len(events) == 1 and not r.has_events()
    assert len(events[0].bunches) == 42629
    # the run end block: 'RUNE', the run number, the number of events
    end = r.end
    assert end[0] == np.frombuffer(b"RUNE", np.float32)[0] and end[1] == 7.0 and end[2] == 1.0
    from merlict_development_kit_amd._native import MlgpuError
    with pytest.raises(MlgpuError):
        r.next_event()


def test_eventio_errors(tmp_path):
    """EventIoTest.cpp:16-34: a damaged sync word and an empty file are
    errors (the reference throws); so is a missing file."""
    from merlict_development_kit_amd._native import MlgpuError
    raw = bytearray(GOLD.read_bytes()[:100])
    raw[2] = 0
    bad = tmp_path / "bad_sync.dat"
    bad.write_bytes(bytes(raw))
    empty = tmp_path / "empty.dat"
    empty.write_bytes(b"")
    for p in (bad, empty, tmp_path / "missing.dat"):
        with pytest.raises(MlgpuError):
            _run(p)
    # a file cut inside the first event's bunches
    cut = tmp_path / "cut.dat"
    cut.write_bytes(GOLD.read_bytes()[:200_000])
    with pytest.raises(MlgpuError):
        _run(cut)


def _first_bunch_block(raw: bytes) -> int:
    """Byte offset of the bunch count of the first photon-bunch object (type
    1205, a sub-object: type/version word, id, length, then array and
    telescope int16, photons float32, bunch count int32), found by its
    self-consistent length."""
    import struct
    for o in range(0, len(raw) - 24, 4):
        word, _, length = struct.unpack_from("<iii", raw, o)
        if word & 0xFFFF != 1205:
            continue
        nb = struct.unpack_from("<i", raw, o + 20)[0]
        elem = 2 if (word >> 20) // 1000 == 1 else 4
        if 0 < nb and length == 12 + 8 * elem * nb:
            return o + 20
    raise AssertionError("no photon-bunch object in the fixture")


def test_eventio_rejects_counts_beyond_their_object(tmp_path):
    """A bunch or offset count that the object's length cannot hold is a
    format error before anything is allocated (a damaged count of 2^31 - 1
    bunches would otherwise ask for 64 GiB)."""
    import struct
    from merlict_development_kit_amd._native import MlgpuError
    raw = bytearray(GOLD.read_bytes())
    at = _first_bunch_block(bytes(raw))
    for count in (2 ** 31 - 1, struct.unpack_from("<i", raw, at)[0] + 1):
        bad = bytearray(raw)
        struct.pack_into("<i", bad, at, count)
        path = tmp_path / f"count_{count}.dat"
        path.write_bytes(bytes(bad))
        with pytest.raises(MlgpuError, match="longer than their object"):
            run = _run(path)
            while run.has_events():
                run.next_event()


def test_eventio_matches_reference_reader():
    """Every bunch of every event, in order, equals what the reference's
    eventio::Run reads (bit for bit: the compact int16 bunches scaled by
    the same float <- double steps, the emission height through the same
    pow)."""
    if not O.ref_available():
        pytest.skip("reference harness not built")
    want = O.ref_corsika_extract(GOLD)
    got = np.concatenate([e.bunches for e in _run()])
    assert got.shape == want.shape
    np.testing.assert_array_equal(got.view(np.uint32), np.ascontiguousarray(want, np.float32).view(np.uint32))


# ------------------------------------------------------------ event tapes
TAPE = Path(__file__).parent / "golden" / "corsika_event_tape.tar"


def _tape(path=TAPE):
    from merlict_development_kit_amd.corsika import EventTapeRun
    return EventTapeRun(path)


def _tar_members(path):
    """The tape's members as Python's tarfile reads them: name -> bytes."""
    import tarfile
    with tarfile.open(path) as tf:
        return {m.name: tf.extractfile(m).read() for m in tf.getmembers() if m.isfile()}


def test_event_tape_read_one_by_one():
    """EventTapeTest.cpp:10-27: RUNH, then three events whose headers start
    with 'EVTH', and no fourth."""
    r = _tape()
    assert r.header[0] == np.frombuffer(b"RUNH", np.float32)[0]
    evth = np.frombuffer(b"EVTH", np.float32)[0]
    for k in range(3):
        assert r.has_events()
        assert r.next_event().header[0] == evth
    assert not r.has_events()
    from merlict_development_kit_amd._native import MlgpuError
    with pytest.raises(MlgpuError):
        r.next_event()


def test_event_tape_matches_tar_members():
    """Headers and bunches bit for bit as Python's tarfile reads the
    members (RUNH, each EVTH, each event's cer.x8.float32 blocks in order)."""
    m = _tar_members(TAPE)
    r = _tape()
    assert r.header.tobytes() == m["000000001/RUNH.float32"]
    for k, ev in enumerate(r, start=1):
        d = f"000000001/{k:09d}"
        assert ev.header.tobytes() == m[f"{d}/EVTH.float32"]
        blocks = sorted(n for n in m if n.startswith(d + "/") and n.endswith(".cer.x8.float32"))
        assert ev.bunches.tobytes() == b"".join(m[n] for n in blocks) and len(ev.bunches) > 0
    assert k == 3


def test_event_tape_photons_point_down():
    """EventTapeTest.cpp:29-56: the photons of every event's bunches (the
    oracle's EventIoPhotonFactory restatement) have wavelengths within
    190-700 nm and directions down, steeper than |z| > 0.7."""
    for ev in _tape():
        ph, _ = O.corsika_photons(np.ascontiguousarray(ev.bunches), seed=0, id_offset=0)
        assert len(ph["wavelength"]) > 0
        assert np.all((ph["wavelength"] > 190e-9) & (ph["wavelength"] < 700e-9))
        assert np.all(ph["direction_z"] < 0) and np.all(np.abs(ph["direction_z"]) > 0.7)


def test_event_tape_errors(tmp_path):
    """A tape that does not start with ddddddddd/RUNH.float32, a damaged
    first header (bad checksum) and an empty file are errors; a tape cut
    after its first event ends the run after it."""
    import tarfile
    import io
    from merlict_development_kit_amd._native import MlgpuError
    m = _tar_members(TAPE)
    wrong = tmp_path / "wrong.tar"
    with tarfile.open(wrong, "w", format=tarfile.USTAR_FORMAT) as tf:
        ti = tarfile.TarInfo("000000001/EVTH.float32")
        ti.size = 1092
        tf.addfile(ti, io.BytesIO(m["000000001/000000001/EVTH.float32"]))
    raw = bytearray(TAPE.read_bytes())
    raw[0] ^= 1  # the name changes, the stored checksum does not
    bad = tmp_path / "bad.tar"
    bad.write_bytes(bytes(raw))
    empty = tmp_path / "empty.tar"
    empty.write_bytes(b"")
    for p in (wrong, bad, empty, tmp_path / "missing.tar"):
        with pytest.raises(MlgpuError):
            _tape(p)
    one = tmp_path / "one.tar"
    with tarfile.open(one, "w", format=tarfile.USTAR_FORMAT) as tf:
        for name in ("000000001/RUNH.float32", "000000001/000000001/EVTH.float32",
                     "000000001/000000001/000000001.cer.x8.float32"):
            ti = tarfile.TarInfo(name)
            ti.size = len(m[name])
            tf.addfile(ti, io.BytesIO(m[name]))
    r = _tape(one)
    events = list(r)
    assert len(events) == 1 and events[0].bunches.tobytes() == m["000000001/000000001/000000001.cer.x8.float32"]
