"""The reference-side binding merlict::gpu (merlict_development_kit_amd/facade)
against the reference's own propagation.

oracle/_ref/facade_check (built by `make -C oracle ref facade` where the
reference sources are) drives both on the same photons: the binding through
merlict::gpu::propagate_photons_in_frame_with_config with a Philox generator,
the reference through merlict::Propagator with PhiloxGenerator(key, i), key =
the seed the binding draws with create_seed(); then sensor::Sensors::
assign_photons on both vectors.

Bars per photon: history length, every intersected object and side, and the
final interaction equal, except boundary flips at a rate below FLIP_RATE;
distances and time of flight within 1e-9 relative, positions, normals and
arrival x, y, theta within 1e-9 (FP64 throughout).

The binding's own symbols and its refusal of a scene it does not model run
without a GPU.
"""
import json
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import pyoracle as O
import photon_sets as sources
from merlict_development_kit_amd import SceneDescription

ROOT = Path(__file__).resolve().parent.parent
CHECK = ROOT / "oracle" / "_ref" / "facade_check"
BINDING = ROOT / "oracle" / "_ref" / "libmerlict_gpu.so"

FLIP_RATE = 1e-5
RTOL = 1e-9
ATOL = 1e-9

needs_check = pytest.mark.skipif(not CHECK.exists(), reason="oracle/_ref/facade_check not built (make -C oracle facade)")


def run_check(*args, timeout=600) -> dict:
    print("facade_check", *args, flush=True)
    proc = subprocess.run([str(CHECK), *map(str, args)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                          timeout=timeout)
    assert proc.returncode == 0, proc.stderr
    out = json.loads(proc.stdout.strip().splitlines()[-1])
    print(out, flush=True)
    return out


@needs_check
def test_binding_exports_reference_signatures():
    nm = subprocess.run(["nm", "-DC", "--defined-only", str(BINDING)], stdout=subprocess.PIPE, text=True, check=True)
    sig = ("(std::vector<merlict::Photon, std::allocator<merlict::Photon> >*, merlict::Frame const*, "
           "merlict::PropagationConfig const*, merlict::random::Generator*)")
    for name in ("propagate_photons_in_frame_with_config", "propagate_photons_in_frame_with_config_multi_thread"):
        assert f"merlict::gpu::{name}{sig}" in nm.stdout, name


@needs_check
def test_binding_rejects_unmodelled_surface_before_any_gpu_work():
    out = run_check("unsupported")
    assert out["gpu"] == "invalid_argument"
    assert out["names_frame"]
    assert out["num_interactions"] == 1  # the photon is untouched


def check_bars(out: dict):
    n = out["photons"]
    flips = out["history_length_differs"] + out["object_differs"] + out["interaction_differs"]
    assert flips <= max(1, int(FLIP_RATE * n)), out
    assert out["reference_failed"] == 0
    assert out["max_rel_distance"] <= RTOL
    assert out["max_rel_time_of_flight"] <= RTOL
    for k in ("max_position", "max_normal", "max_support", "max_direction", "max_incident", "max_arrival_xy",
              "max_arrival_txy"):
        assert out[k] <= ATOL, k
    assert out["max_rel_arrival_time"] <= RTOL
    assert out["arrivals_gpu"] > 0
    assert abs(out["arrivals_gpu"] - out["arrivals_ref"]) <= flips
    assert out["arrivals_mismatched"] <= 2 * flips


@pytest.mark.gpu
@needs_check
@pytest.mark.parametrize("scene,n,max_int", [("lens", 50_000, 5), ("fact_lg", 50_000, 8)])
def test_binding_matches_reference_on_scene_files(scene, n, max_int):
    desc = SceneDescription.lens() if scene == "lens" else SceneDescription.fact(light_guides=True)
    photons = (sources.lens_beam(n, seed=31) if scene == "lens" else sources.fact_bunch(n, seed=32)).numpy()
    with tempfile.TemporaryDirectory() as d:
        sp, pp = Path(d) / "scene.bin", Path(d) / "photons.bin"
        desc.save(str(sp))
        O.write_photons(pp, photons)
        out = run_check("compare", sp, pp, 4242, max_int, 16)
    assert out["photons"] == n
    check_bars(out)


@pytest.mark.gpu
@needs_check
def test_binding_matches_reference_on_api_built_scene():
    # BiConvexLens with restricted propagation, RectangularBox, Cylinder from
    # end points, Triangle from vertices, axis-angle poses, 49 sensors
    out = run_check("api", 50_000, 77, 16)
    check_bars(out)


@pytest.mark.gpu
@needs_check
def test_binding_raises_out_of_range_like_func1():
    out = run_check("errors")
    assert out["reference"] == "out_of_range"
    assert out["gpu"] == "out_of_range"
    k = out["num_interactions"]
    assert all(v > 1 for v in k[:5])  # propagated before the failing photon
    assert all(v == 1 for v in k[6:])  # untouched after it
