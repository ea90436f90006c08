"""Solvent mixtures (per-site densities; no counterpart in the reference,
whose solvent is one molecule at one density, hnc3d.c:90-92): the 1D XRISM
of oracle/rism1d.py rism_vv_mixture, the 3D solute on its chi
(oracle/rism3d.py Solute3D(site_rho=)), and the C++ producer
rism3d_solvent_chi_mixture.

Pinned by an invariant that holds only with the densities and the
transposed contraction in the right places: water split into two copies at
0.3 rho and 0.7 rho is the same liquid, so every copy-copy block of t
equals pure water's t, and a solute's mu equals its mu in pure water.  A
density on the wrong index (rho_j for rho_i, or chi_ij for chi_ji in the
solute contraction or in tau) would weight the copies as 0.6 / 1.4 instead
and move both."""
import numpy as np
import pytest

from oracle import closures as C
from oracle import rism1d as R1
from oracle import rism3d as O

BETA, RHO = 1.6889, 0.0333295
NRAD, RMAX = 512, 40.0
NA = ("Na", (0.0, 0.0, 0.0), 2.35, 0.13, 1.0)
CL = ("Cl", (0.0, 0.0, 0.0), 4.40, 0.10, -1.0)
RHO_ION = 0.1 * 6.02214076e-4  # 0.1 M in A^-3


def _water():
    from bgy3d_amd.solutes import find_molecule
    return find_molecule("water, PR-SPC/E")


@pytest.fixture(scope="module")
def copies():
    W = _water()
    pure = R1.rism_vv(C.KH, NRAD, RMAX, BETA, RHO, W, eps=0.0)
    rho = [0.3 * RHO] * 3 + [0.7 * RHO] * 3
    mix = R1.rism_vv_mixture(C.KH, NRAD, RMAX, BETA, list(W) + list(W), [0, 0, 0, 1, 1, 1], rho)
    return W, pure, mix, rho


def test_copies_1d(copies):
    W, pure, mix, rho = copies
    for A in range(2):
        for B in range(2):
            blk = mix["t"][3 * A:3 * A + 3, 3 * B:3 * B + 3]
            assert np.max(np.abs(blk - pure["t"])) < 1e-8
            # the columns summed over the copies act as pure water's chi
    chi_sum = mix["chi"][:3, :3] + mix["chi"][3:, :3]  # sum over B of chi_{jB, iA=0}
    assert np.max(np.abs(chi_sum - pure["chi"])) < 1e-8
    # and not when the density sits on the column index
    assert np.max(np.abs(mix["chi"][:3, :3] + mix["chi"][:3, 3:] - pure["chi"])) > 1e-3


def test_copies_3d(copies):
    W, pure, mix, rho = copies
    from bgy3d_amd.solutes import find_molecule
    sol = find_molecule("water, cSPC/E")
    pd = O.Problem.from_cli(8.0, 32, BETA, RHO, closure=C.KH, rmax=RMAX, nrad=NRAD)
    a = O.Solute3D(pd, W, sol, pure["chi"])
    ta, _ = a.solve(atol=1e-11)
    b = O.Solute3D(pd, list(W) + list(W), sol, mix["chi"], site_rho=rho)
    tb, _ = b.solve(atol=1e-11)
    ma, mb = a.observables(ta), b.observables(tb)
    assert abs(ma["mu"] - mb["mu"]) < 1e-8
    assert abs(ma["pmv_factor"] - mb["pmv_factor"]) < 1e-8
    assert np.max(np.abs(mb["g"][:3] - ma["g"])) < 1e-8
    assert np.max(np.abs(mb["g"][3:] - ma["g"])) < 1e-8


def test_saline_producer_matches_oracle():
    """rism3d_solvent_chi_mixture (C++) against rism_vv_mixture: water with
    Na+ / Cl- at 0.1 M, KH."""
    from bgy3d_amd import rism3d as R
    W = _water()
    sites = list(W) + [NA, CL]
    mol = [0, 0, 0, 1, 2]
    rho = [RHO] * 3 + [RHO_ION] * 2
    ref = R1.rism_vv_mixture(C.KH, NRAD, RMAX, BETA, sites, mol, rho)
    out = R.rism_solvent_mixture(sites, mol, rho, {"rmax": RMAX, "nrad": NRAD, "beta": BETA,
                                                   "rho": RHO, "closure": "KH", "L": 8, "N": 32})
    chi = out["susceptibility"]
    assert np.max(np.abs(chi - ref["chi"])) < 1e-8
    # water-water block near pure water's (the ions are dilute), and chi is
    # not symmetric: rho_i h_ij
    pure = R1.rism_vv(C.KH, NRAD, RMAX, BETA, RHO, W, eps=0.0)
    assert np.max(np.abs(chi[:3, :3] - pure["chi"])) < 0.05
    assert np.max(np.abs(chi[3, 0] - chi[0, 3])) > 1e-3


def test_mixture_validation():
    from bgy3d_amd import rism3d as R
    W = _water()
    s = {"rmax": RMAX, "nrad": NRAD, "beta": BETA, "rho": RHO, "closure": "KH", "L": 8, "N": 32}
    with pytest.raises(R.RismError):  # one molecule, two densities
        R.rism_solvent_mixture(list(W), [0, 0, 0], [RHO, RHO, 0.5 * RHO], s)
    with pytest.raises(ValueError):
        R.rism_solvent_mixture(list(W), [0, 0], [RHO] * 3, s)
