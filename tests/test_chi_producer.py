"""The product's 1D DRISM chi producer (bgy3d_amd/csrc/rism1d.cpp, host C++
behind rism3d_solvent_chi) against the oracle's chi (tests/golden, pinned by
README:9-16 and README:20-31).  Runs on the CPU: no GPU is involved."""
import numpy as np
import pytest

from conftest import CHI_CSPCE_KH, CHI_W, load_chi

CASES = [
    ("water, PR-SPC/E", "HNC", 40.0, 1536, CHI_W),
    ("water, PR-SPC/E", "HNC", 160.0, 4096, "chi_prspce_hnc_eps78.4_rmax160_nrad4096.npy"),
    ("water, cSPC/E", "KH", 160.0, 4096, CHI_CSPCE_KH),
]


@pytest.mark.parametrize("solvent,closure,rmax,nrad,fixture", CASES)
def test_chi_matches_oracle(solvent, closure, rmax, nrad, fixture):
    from bgy3d_amd import rism3d as R
    chi = R.rism_solvent(solvent, {"rmax": rmax, "nrad": nrad, "rho": 0.0333295, "beta": 1.6889,
                                   "dielectric": 78.4, "closure": closure})["susceptibility"]
    np.testing.assert_allclose(chi, load_chi(fixture), atol=1e-10)


def test_chi_symmetric_and_plain_rism():
    from bgy3d_amd import rism3d as R
    chi = R.rism_solvent("water, PR-SPC/E", {"rmax": 20.0, "nrad": 256, "rho": 0.0333295,
                                             "beta": 1.6889})["susceptibility"]
    np.testing.assert_allclose(chi, np.transpose(chi, (1, 0, 2)), atol=1e-10)
    # chi -> omega as k -> large: the intramolecular bond sinc plus a vanishing rho h
    assert abs(chi[0, 0, -1] - 1.0) < 1e-2
