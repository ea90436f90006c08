"""Writes config0_water_96.json: the oracle's plain FP64 HNC solve of the
PR-SPC/E water solute in PR-SPC/E water with chi from the 1D run at
rmax 160 A, nrad 4096 (chi_prspce_hnc_eps78.4_rmax160_nrad4096.npy),
--L 10 --N 96, rho 0.0333295, beta 1.6889, renormalised (tau) path, to
max|F| <= 1e-13.  Stored: mu, and per solvent site the sum of g over the
grid, max g and g at three grid points, so that a GPU solve of the same
workload can be checked without a 21 MB array.

    python tests/golden/make_config0.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import rism3d  # noqa: E402
from bgy3d_amd.solutes import find_molecule  # noqa: E402

POINTS = [(48, 48, 48), (40, 52, 47), (10, 70, 33)]  # (z, y, x) grid indices


def main():
    chi = np.load(os.path.join(HERE, "chi_prspce_hnc_eps78.4_rmax160_nrad4096.npy"))
    w = find_molecule("water, PR-SPC/E")
    pd = rism3d.Problem.from_cli(10.0, 96, 1.6889, 0.0333295, rmax=160.0, nrad=4096)
    out = rism3d.solute_solve(pd, w, w, chi, atol=1e-13)
    g = out["g"]
    res = {"settings": {"N": 96, "L": 10.0, "rho": 0.0333295, "beta": 1.6889, "rmax": 160.0,
                        "nrad": 4096, "closure": "HNC", "atol": 1e-13},
           "mu": float(out["mu"]),
           "points": POINTS,
           "sites": [{"sum": float(g[i].sum()), "max": float(g[i].max()),
                      "at": [float(g[i][p]) for p in POINTS]} for i in range(g.shape[0])]}
    with open(os.path.join(HERE, "config0_water_96.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps({"mu": res["mu"], "iters": str(out["stats"])[:200]}))


if __name__ == "__main__":
    main()
