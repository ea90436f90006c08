"""The 3D solvent-solvent solve (oracle/solvent3d.py, hnc3d_solvent_solve
and compute_t2_m, hnc3d.c:231-324) for more than three sites: methanol,
six sites and 21 site-site pairs.

The reference has no regression output for such a solvent.  The oracle's
per-k OZ solve, X = (1 - rho W C)^-1 W C W, is general in m and pinned at
m = 3 by test/out/water.hnc3d.g2 (test_oracle_solvent3d.py).  Here its
indexing is pinned by a symmetry: listing the sites of the molecule in
another order must permute the pair functions and leave mu unchanged, to
the solve's tolerance.  A wrong pair index p(i, j) for i >= 3, a transposed
omega or C, or a pair missed in the 6 x 6 matrices breaks this.  The two
solves run in two worker processes (one BLAS thread each) at 24^3.
"""
import multiprocessing as mp
import os

import numpy as np
import pytest

BETA, RHO, N, L, RMAX, NRAD = 1.6889, 0.0149, 24, 12.0, 40.0, 512
PERM = (4, 5, 0, 1, 2, 3)  # O and OH first: new site k is old site PERM[k]


def _solve(order):
    from oracle import closures as C
    from oracle import rism1d as R1
    from oracle import rism3d as O
    from oracle import solvent3d as S3
    from bgy3d_amd.solutes import find_molecule
    meoh = find_molecule("methanol")
    mol = tuple(meoh[k] for k in order)
    pd = O.Problem.from_cli(L, N, BETA, RHO, closure=C.KH, rmax=RMAX, nrad=NRAD)
    s = S3.Solvent3D(pd, mol)
    one = R1.rism_vv(C.KH, NRAD, RMAX, BETA, RHO, mol)
    t, it = s.solve(t0=s.initial_guess(one["t"], RMAX), atol=1e-10, max_iter=2000)
    obs = s.observables(t)
    return {"pairs": s.pairs, "g": obs["g"], "it": it,
            "mu": (obs["mu_hnc"], obs["mu_kh"], obs["mu_gf"])}


@pytest.fixture(scope="module")
def runs():
    keys = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
    saved = {k: os.environ.get(k) for k in keys}
    for k in keys:
        os.environ[k] = "1"
    try:
        with mp.get_context("spawn").Pool(2) as pool:
            a, b = pool.map(_solve, [tuple(range(6)), PERM])
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return a, b


def test_methanol_mu_invariant_under_site_order(runs):
    a, b = runs
    print(f"  mu {a['mu']} reordered {b['mu']} iterations {a['it']} {b['it']}")
    assert np.all(np.isfinite(a["mu"]))
    np.testing.assert_allclose(b["mu"], a["mu"], rtol=0, atol=1e-8)


def test_methanol_pairs_permute_with_the_sites(runs):
    a, b = runs
    index = {ij: p for p, ij in enumerate(a["pairs"])}
    worst = 0.0
    for q, (i, j) in enumerate(b["pairs"]):
        oi, oj = PERM[i], PERM[j]
        p = index[(max(oi, oj), min(oi, oj))]
        worst = max(worst, float(np.max(np.abs(b["g"][q] - a["g"][p]))))
    print(f"  max |g_perm - g| {worst:.2e}")
    assert worst < 1e-8
