"""GPU photon sources (mlgpu_generate_photons) against the oracle's
restatement: every photon bit for bit, since both use only correctly rounded
IEEE operations on the same Philox draws."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import pyoracle as O
from merlict_development_kit_amd import sources as S
from merlict_development_kit_amd._native import MlgpuError

pytestmark = pytest.mark.gpu

SOURCES = {
    "lens_beam": S.lens_beam(),
    "fact_bunch": S.fact_bunch(),
    "calibration": S.calibration(),
    "point_like": S.point_like(math.radians(10.0)),
    "parallel_disc": S.parallel_disc(0.3),
    "posed_bunch": S.fact_bunch().posed(S.rotation_xyz(0.02, -0.03, 1.0), (0.5, -0.25, 0.0)),
}


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.mark.parametrize("name", sorted(SOURCES))
def test_gpu_source_matches_oracle_bitwise(cuda, name):
    src = SOURCES[name]
    n = 100_000
    got = src.generate(n, seed=21, photon_id_offset=5, device=cuda).numpy()
    want = O.generate_photons(src.to_c(O.Source), n, seed=21, photon_id_offset=5)
    for k in want:
        mism = int((got[k] != want[k]).sum())
        assert mism == 0, f"{name}: {mism} of {n} photons differ in {k}"


def test_gpu_source_rejects_bad_parameters(cuda):
    for bad in (S.Source(direction="tilt", angle=2.0), S.Source(lambda_min=0.0, lambda_max=1e-7),
                S.Source(lambda_min=500e-9, lambda_max=400e-9), S.Source(support="disc", radius=-1.0)):
        with pytest.raises(MlgpuError):
            bad.generate(10, device=cuda)
    assert len(S.fact_bunch().generate(0, device=cuda)) == 0
    with pytest.raises(ValueError):
        S.fact_bunch().generate(10, device="cpu")
