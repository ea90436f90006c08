"""Oracle solves of the test/run-ions.scm rows, run in worker processes.

The settings are test/run-ions.scm:16-30: cSPC/E water, rmax 160,
nrad 4096, --L 10 --N 64, rho 0.0333295, beta 1.6889, dielectric 78.4.
"""
import json
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "run_ions_out.json")
SOLVENT = "water, cSPC/E"
BETA, RHO, RMAX, NRAD = 1.6889, 0.0333295, 160.0, 4096


def golden():
    with open(GOLDEN) as f:
        d = json.load(f)
    return {dim: {clo: {ion: float(v) for ion, v in rows.items()} for clo, rows in tab.items()}
            for dim, tab in d.items()}


def solve_3d(args):
    """(closure name, ion, chi (m, m, nrad), drop_k0) -> mu of the 3D oracle
    at N = 64, to max|F| <= 1e-9.  drop_k0 plants a mistake in the oracle:
    the k = 0 term tau(0) F(0) is left out of the renormalisation."""
    clo, ion, chi, drop_k0 = args
    from oracle import closures as C
    from oracle import rism3d
    from bgy3d_amd.solutes import find_molecule
    rism3d.FFT_WORKERS = 2
    pd = rism3d.Problem.from_cli(10.0, 64, BETA, RHO, closure=C.NAMES[clo], rmax=RMAX,
                                 nrad=NRAD)
    s = rism3d.Solute3D(pd, find_molecule(SOLVENT), find_molecule(ion), chi)
    if drop_k0:
        s.tau[:, 0, 0, 0] = 0.0
    t, _ = s.solve(atol=1e-9)
    return s.observables(t)["mu"]
