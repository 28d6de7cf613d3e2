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
    assert N.lib().mlgpu_abi_version() == 1


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


def test_pixel_grid_is_hexagonal_and_unique():
    a = SceneDescription.fact(light_guides=False).arrays()
    pix = a["pos"][a["sensor_id"] >= 0][:, :2]
    assert len(np.unique(np.round(pix, 9), axis=0)) == 1440
    # nearest-neighbour distance is the 9.5 mm pitch
    from scipy.spatial import cKDTree
    d, _ = cKDTree(pix).query(pix, k=2)
    np.testing.assert_allclose(d[:, 1], 9.5e-3, rtol=1e-9)
