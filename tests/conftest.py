import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    return load


def split_golden(g):
    photons = {k[3:]: g[k] for k in g if k.startswith("in_")}
    results = {k[4:]: g[k] for k in g if k.startswith("out_")}
    return photons, results


@pytest.fixture(scope="session")
def oracle_threads():
    return max(1, min(8, os.cpu_count() or 1))
