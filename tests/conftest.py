import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_chi(name):
    return np.load(os.path.join(GOLDEN, name))


CHI_W = "chi_prspce_hnc_eps78.4_rmax40_nrad1536.npy"
CHI_CSPCE_KH = "chi_cspce_kh_eps78.4_rmax160_nrad4096.npy"

# Physics of the reference README 3D example (README:20-31).
README = {"rho": 0.0333295, "beta": 1.6889}


@pytest.fixture(scope="session")
def chi_w():
    return load_chi(CHI_W)
