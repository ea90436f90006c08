"""Writes methanol_solvent3d_32.json: the oracle's 3D solvent-solvent solve
(oracle/solvent3d.py, hnc3d_solvent_solve) of methanol, six sites and 21
site-site pairs, at N = 32, L = 12 A, KH, beta 1.6889, rho 0.0149,
rmax 40, nrad 512, started from the 1D RISM t_ij(r) (oracle/rism1d.py) and
converged to max|F| <= 1e-11: mu by the HNC, KH and GF functionals and, per
pair in the reference's order (i >= j), the charge, dipole and diagonal
second moments of 1 - g_ij.  The reference has no regression output for a
3D solvent of more than three sites; the oracle's per-k OZ solve is
general in m and is pinned by test/out/water.hnc3d.g2 at m = 3 and by its
invariance under a reordering of the sites (tests/test_oracle_solvent3d_m.py).

Run from the repository root:  python tests/golden/make_methanol_solvent3d.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import closures as C  # noqa: E402
from oracle import rism1d as R1  # noqa: E402
from oracle import rism3d as O  # noqa: E402
from oracle import solvent3d as S3  # noqa: E402
from bgy3d_amd.solutes import find_molecule  # noqa: E402

BETA, RHO, N, L, RMAX, NRAD = 1.6889, 0.0149, 32, 12.0, 40.0, 512
OUT = os.path.join(HERE, "methanol_solvent3d_32.json")


def main():
    meoh = find_molecule("methanol")
    pd = O.Problem.from_cli(L, N, BETA, RHO, closure=C.KH, rmax=RMAX, nrad=NRAD)
    s = S3.Solvent3D(pd, meoh)
    one = R1.rism_vv(C.KH, NRAD, RMAX, BETA, RHO, meoh)
    t, it = s.solve(t0=s.initial_guess(one["t"], RMAX), atol=1e-11, max_iter=2000)
    obs = s.observables(t)
    pairs = []
    for p, (i, j) in enumerate(s.pairs):
        q, d, Q = S3.moments(pd, 1.0 - obs["g"][p])
        pairs.append({"i": i, "j": j, "q": q, "d": list(d), "Qdiag": [Q[a][a] for a in range(3)]})
    out = {"settings": {"N": N, "L": L, "beta": BETA, "rho": RHO, "closure": "KH",
                        "rmax": RMAX, "nrad": NRAD, "tol": 1e-11},
           "iterations": it, "mu_hnc": obs["mu_hnc"], "mu_kh": obs["mu_kh"],
           "mu_gf": obs["mu_gf"], "pairs": pairs}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(OUT, it, obs["mu_hnc"], obs["mu_kh"], obs["mu_gf"])


if __name__ == "__main__":
    main()
