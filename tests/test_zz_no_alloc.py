"""mlgpu_propagate_noalloc: the propagation entry point that allocates
nothing (its wavefront workspace is a static device array of the library),
called through the C ABI and checked bit for bit against mlgpu_propagate on
two calls in a row, the second with a smaller batch, and its size limit."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

FIELDS = ("final_frame", "sensor_id", "final_interaction", "num_interactions", "error", "time_of_flight", "x", "y",
          "cx", "cy", "accumulated_distance")


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.mark.gpu
def test_gpu_noalloc_propagation_equals_mlgpu_propagate(cuda):
    from merlict_development_kit_amd import GpuScene, PropagationConfig, PropagationResult, SceneDescription, sources
    from merlict_development_kit_amd import _native as N
    lib = N.lib()
    max_photons = 262_144
    scene = GpuScene(SceneDescription.fact(light_guides=True))
    src = sources.fact_bunch()
    stream = torch.cuda.current_stream()
    calls = [(schedule, n, seed) for schedule in ("wavefront", "persistent")
             for n, seed in ((max_photons, 5), (100_000, 6))]
    for call, (schedule, n, seed) in enumerate(calls):
        scene.set_tuning(schedule=schedule)
        print(f"call {call}: {schedule}, {n} photons, seed {seed}", flush=True)
        ph = src.generate(n, seed=seed, photon_id_offset=1000 * call, device=cuda)
        cfg = PropagationConfig(5, seed=seed + 10, photon_id_offset=77).c_struct()
        want, got = PropagationResult.empty(n, cuda), PropagationResult.empty(n, cuda)
        for r in (want, got):
            for f in FIELDS:
                getattr(r, f).fill_(0x5A if getattr(r, f).dtype == torch.uint8 else 123)
        ps = ph.c_struct()
        ws, gs = want.c_struct(), got.c_struct()
        st = C.c_void_p(stream.cuda_stream)
        assert lib.mlgpu_propagate(scene.handle, 0, C.byref(ps), n, C.byref(cfg), C.byref(ws), st) == 0
        assert lib.mlgpu_propagate_noalloc(scene.handle, 0, C.byref(ps), n, C.byref(cfg), C.byref(gs), st) == 0, \
            lib.mlgpu_last_error()
        torch.cuda.synchronize()
        for f in FIELDS:
            a, b = getattr(want, f), getattr(got, f)
            assert torch.equal(a, b), f
        det = int((got.sensor_id >= 0).sum())
        print(f"call {call}: equal, {det} detected", flush=True)
        assert det > 0.3 * n
    # beyond the limit: an argument error, nothing launched
    n = max_photons + 1
    ph = src.generate(n, seed=1, device=cuda)
    out = PropagationResult.empty(n, cuda)
    ps, rs = ph.c_struct(), out.c_struct()
    cfg = PropagationConfig(5).c_struct()
    assert lib.mlgpu_propagate_noalloc(scene.handle, 0, C.byref(ps), n, C.byref(cfg), C.byref(rs), None) == 1
    assert b"262144" in lib.mlgpu_last_error()
