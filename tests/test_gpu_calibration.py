"""Plenoscope light-field calibration on the GPU (mlgpu_calibration_photons,
mlgpu_lixel_filler_*) against the reference's LixelStatisticsFiller (golden
bytes of calibration::write) and the oracle."""
import math

import numpy as np
import pytest
import torch

import pyoracle as O
from merlict_development_kit_amd import GpuScene, PropagationConfig, SceneDescription
from merlict_development_kit_amd import calibration as CAL
from merlict_development_kit_amd._native import MlgpuError

pytestmark = pytest.mark.gpu

PLENOSCOPE_LIXELS = 515_023  # lixels of the default plenoscope's light-field sensor


@pytest.fixture(scope="module")
def cuda():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _words(rec: np.ndarray, dev) -> torch.Tensor:
    return torch.tensor(np.ascontiguousarray(rec, CAL.CALIBRATION_PHOTON).view(np.int32), device=dev)


def test_gpu_lixel_filler_matches_reference(cuda, golden):
    """Three blocks of records into the GPU filler: the statistics are the
    bytes calibration::write wrote for the reference's filler (NaNs
    included)."""
    g = golden("lixel_statistics")
    rec = g["records"].view(CAL.CALIBRATION_PHOTON)
    nl, nb, pb = (int(v) for v in g["config"])
    f = CAL.LixelStatisticsFiller(nl, nb, pb)
    for b in range(nb):
        f.fill_in_block(_words(rec[b * pb:(b + 1) * pb], cuda))
    st = f.get_lixel_statistics()
    f.close()
    assert st.tobytes() == g["written"].tobytes()


def test_gpu_plenoscope_calibration_matches_oracle(cuda):
    """Two blocks of 2^19 calibration photons through the plenoscope (515,023
    lixels): the GPU's records equal the oracle's for the same results, and
    the statistics after both blocks equal the oracle's filler byte for
    byte."""
    desc = SceneDescription.plenoscope()
    gs = GpuScene(desc)
    src = CAL.calibration_source(35.5, math.radians(3.25))
    n = 1 << 19
    filler = CAL.LixelStatisticsFiller(PLENOSCOPE_LIXELS, 2, n)
    all_rec = []
    for b in range(2):
        ph = src.generate(n, seed=11, photon_id_offset=b * n, device=cuda)
        res = gs.propagate(ph, PropagationConfig(5, seed=11, photon_id_offset=b * n))
        words = CAL.calibration_photons(src, 11, b * n, res, PLENOSCOPE_LIXELS)
        rec = words.cpu().numpy().view(CAL.CALIBRATION_PHOTON)
        orec = O.calibration_photons(src.to_c(O.Source), 11, b * n, res.numpy(), PLENOSCOPE_LIXELS)
        assert rec.tobytes() == orec.tobytes()
        filler.fill_in_block(words)
        all_rec.append(orec)
    st = filler.get_lixel_statistics()
    filler.close()
    ost = O.lixel_statistics(np.concatenate(all_rec), PLENOSCOPE_LIXELS, 2, n)
    assert st.tobytes() == ost.tobytes()
    reached = np.concatenate(all_rec)["reached_sensor"].sum()
    assert reached > 0.3 * 2 * n
    # the efficiencies add up to the reached fraction times the lixels
    assert float(st["efficiency"].astype(np.float64).sum()) == pytest.approx(
        reached / (2 * n) * PLENOSCOPE_LIXELS, rel=1e-5)
    seen = st["efficiency"] > 0
    # a lixel sees directions inside the calibration cone
    assert (np.hypot(st["cx_mean"][seen], st["cy_mean"][seen]) < 1.2 * math.radians(3.25)).all()


def test_gpu_run_calibration_and_file(cuda, tmp_path):
    """run_calibration (Calibrator.cpp:140-163) on two blocks; the written
    file reads back as the same records and efficiencies."""
    gs = GpuScene(SceneDescription.plenoscope())
    st = CAL.run_calibration(gs, PLENOSCOPE_LIXELS, 35.5, math.radians(3.25), CAL.CalibrationConfig(2, 1 << 18),
                             seed=4)
    assert len(st) == PLENOSCOPE_LIXELS
    p = tmp_path / "lixel_statistics.bin"
    CAL.write(st, p)
    assert p.stat().st_size == 48 * PLENOSCOPE_LIXELS
    assert CAL.read(p).tobytes() == st.tobytes()
    np.testing.assert_array_equal(CAL.read_efficiencies(p), st["efficiency"])


def test_gpu_lixel_filler_errors(cuda):
    f = CAL.LixelStatisticsFiller(4, 1, 2)
    rec = np.zeros(2, CAL.CALIBRATION_PHOTON)
    rec["reached_sensor"] = 1
    rec["lixel_id"] = [1, 4]
    with pytest.raises(MlgpuError):  # lixel 4 of 4
        f.fill_in_block(_words(rec, cuda))
    f.fill_in_block(torch.zeros(0, dtype=torch.int32, device=cuda))
    st = f.get_lixel_statistics()
    assert (st["efficiency"] == 0).all()  # the bad block added nothing
    f.close()
    with pytest.raises(MlgpuError):
        CAL.LixelStatisticsFiller(0, 1, 1)
