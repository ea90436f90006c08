"""Regenerate the committed fixtures under tests/golden/.

    python tests/golden/make_fixtures.py

The fixtures are solvent susceptibilities chi(k) [m][m][nrad] produced by
the oracle's 1D DRISM restatement (oracle/rism1d.py, rism.f90:310-583), used
as *inputs* to the 3D tests so that GPU tests do not spend minutes in the 1D
solver.  The oracle itself is pinned against the reference README's 1D
anchor (-7.869294) and its 3D anchors, which take this very chi as input
(tests/test_oracle_anchors.py).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import rism1d  # noqa: E402
from bgy3d_amd.solutes import find_molecule  # noqa: E402

FIXTURES = {
    # README:20-31: energy --hnc, PR-SPC/E in PR-SPC/E, beta 1.6889,
    # rho 0.0333295, dielectric 78.4; chi on the grid hnc3d_solute_solve
    # builds itself for --L 10 --N 96 (upscale, hnc3d.c:772-783).
    "chi_prspce_hnc_eps78.4_rmax40_nrad1536.npy": dict(
        solvent="water, PR-SPC/E", closure=0, rmax=40.0, nrad=1536,
        beta=1.6889, rho=0.0333295, eps=78.4),
    # The solvent of the benchmark workload: PR-SPC/E water, HNC, 1D grid
    # rmax 160 A, nrad 4096 (the 1D run of README:9-16).
    "chi_prspce_hnc_eps78.4_rmax160_nrad4096.npy": dict(
        solvent="water, PR-SPC/E", closure=0, rmax=160.0, nrad=4096,
        beta=1.6889, rho=0.0333295, eps=78.4),
    # test/run-ions.scm:16-30 settings, KH closure, cSPC/E water.
    "chi_cspce_kh_eps78.4_rmax160_nrad4096.npy": dict(
        solvent="water, cSPC/E", closure=1, rmax=160.0, nrad=4096,
        beta=1.6889, rho=0.0333295, eps=78.4),
}


def main():
    for name, p in FIXTURES.items():
        vv = rism1d.rism_vv(p["closure"], p["nrad"], p["rmax"], p["beta"], p["rho"],
                            find_molecule(p["solvent"]), eps=p["eps"])
        np.save(os.path.join(HERE, name), vv["chi"])
        print(name, vv["stats"]["iters"], vv["mu"])


if __name__ == "__main__":
    main()
