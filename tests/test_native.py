"""The C ABI library: loads, exports every declared entry point, host-side
functions behave (no GPU compute here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from merlict_development_kit_amd import SceneDescription
from merlict_development_kit_amd import _native as N

HEADER = Path(__file__).resolve().parent.parent / "include" / "mlgpu.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mlgpu_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_function_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    lib = N.lib()
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, f"{name} has no ctypes signature"


def test_abi_version():
    assert N.lib().mlgpu_abi_version() == 2


def test_desc_roundtrip(tmp_path):
    d = SceneDescription.fact(light_guides=True)
    p = tmp_path / "fact.bin"
    d.save(str(p))
    e = SceneDescription.load(str(p))
    a, b = d.arrays(), e.arrays()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_bad_file_is_rejected(tmp_path):
    p = tmp_path / "junk.bin"
    p.write_bytes(b"not a scene at all")
    with pytest.raises(N.MlgpuError):
        SceneDescription.load(str(p))


def test_builder_argument_checks():
    h = C.c_void_p()
    assert N.lib().mlgpu_build_lens_scene(-1.0, 0.25, 1.49, 0.06, C.byref(h)) == 1
    assert b"bad argument" in N.lib().mlgpu_last_error()


def test_scene_sizes():
    assert SceneDescription.lens().num_frames == 24          # root, lens, 2 caps, plane, 19 cells
    fact = SceneDescription.fact(light_guides=False)
    a = fact.arrays()
    assert int((a["kind"] == 1).sum()) == 30                  # facets
    assert int((a["sensor_id"] >= 0).sum()) == 1440           # pixels
    lg = SceneDescription.fact(light_guides=True).arrays()
    assert int((lg["kind"] == 10).sum()) == 1440 * 12          # two triangles per guide wall


def test_fact_pixels_are_the_fact_camera_table():
    """Sensor id = FACT chid at the FACT propagation app's pixel position
    (fact_propagation.cpp:316-341: float millimetres / 1e3), spot-checked
    against values of the app's table; positions unique, neighbours 9.19 to
    9.5 mm apart (the camera's modules sit off a regular lattice)."""
    for lg in (False, True):
        a = SceneDescription.fact(light_guides=lg).arrays()
        frames = np.nonzero(a["sensor_id"] >= 0)[0]
        order = frames[np.argsort(a["sensor_id"][frames])]
        holder = a["mother"][order] if lg else order
        pix = a["pos"][holder][:, :2]
        assert len(pix) == 1440 and np.array_equal(a["sensor_id"][order], np.arange(1440))
        for chid, (x, y) in {0: (28.5, 172.805), 9: (-0.0, 172.805), 36: (114.0, 123.805),
                             1439: (-4.75, 164.54)}.items():
            assert pix[chid, 0] == float(np.float32(x)) / 1e3 and pix[chid, 1] == float(np.float32(y)) / 1e3, chid
        assert len(np.unique(np.round(pix, 9), axis=0)) == 1440
        from scipy.spatial import cKDTree
        d, _ = cKDTree(pix).query(pix, k=2)
        assert 9.18e-3 < d[:, 1].min() and d[:, 1].max() <= 9.5e-3 + 1e-12


def test_fact_tables_match_the_reference_source():
    """fact_tables.inc is what gen_fact_tables.py writes from the reference's
    fact_propagation.cpp (skipped where the reference sources are absent)."""
    import subprocess
    import sys
    src = Path("/root/reference/merlict_fact/apps/fact_propagation.cpp")
    if not src.exists():
        pytest.skip("reference sources absent")
    csrc = Path(__file__).resolve().parent.parent / "merlict_development_kit_amd" / "csrc"
    before = (csrc / "fact_tables.inc").read_text()
    subprocess.run([sys.executable, str(csrc / "gen_fact_tables.py"), str(src)], check=True, capture_output=True)
    assert (csrc / "fact_tables.inc").read_text() == before
