import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "cylinder-fem-amr_amd"
GOLDEN = Path(__file__).resolve().parent / "golden"
for p in (PKG, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def read_csv(path):
    import numpy as np
    return np.genfromtxt(path, delimiter=",", names=True)


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
