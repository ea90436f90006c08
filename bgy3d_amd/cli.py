"""`energy --hnc` of the reference driver (guile/bgy3d.scm:1007-1020).

    python -m bgy3d_amd.cli energy --hnc --norm-tol 1e-14 --dielectric 78.4 \\
        --rho 0.0333295 --beta 1.6889 --L 10 --N 96 \\
        --solvent "water, PR-SPC/E" --solute "water, PR-SPC/E"

As make-pes does (guile/bgy3d.scm:819-861), chi comes from the 1D DRISM
solvent run on the radial grid of the settings (rmax = L, nrad = N unless
--rmax/--nrad are given), then the 3D solve runs on the GPU and mu is
printed.  With --chi-upscale chi is instead built by the 3D solver itself on
its upscaled grid (rmax = 2 max(L), nrad = 16 max(N); hnc3d.c:772-783), the
path that gives README's -6.349052 for the command above.
"""
import argparse
import sys

from . import rism3d as R


def parse(argv):
    # option names of guile/bgy3d.scm:635-688 that this path reads
    ap = argparse.ArgumentParser(prog="bgy3d_amd.cli")
    ap.add_argument("command", choices=["energy"])
    ap.add_argument("--hnc", action="store_true")
    ap.add_argument("--solvent", required=True)
    ap.add_argument("--solute")
    ap.add_argument("--L", type=float, default=10.0)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--rho", type=float, default=0.033427745)
    ap.add_argument("--beta", type=float, default=1.6889)
    ap.add_argument("--norm-tol", type=float, default=1e-7)
    ap.add_argument("--max-iter", type=int, default=1500)
    ap.add_argument("--dielectric", type=float, default=0.0)
    ap.add_argument("--closure", default="HNC")
    ap.add_argument("--rmax", type=float)
    ap.add_argument("--nrad", type=int)
    ap.add_argument("--no-renorm", action="store_true")
    ap.add_argument("--snes-solver", choices=["mdiis", "picard", "newton", "jager", "trial"],
                    default="mdiis")
    ap.add_argument("--lambda", dest="lam", type=float, default=0.02)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--chi-upscale", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args(argv)


def main(argv=None):
    a = parse(sys.argv[1:] if argv is None else argv)
    if not a.hnc:
        print("bgy3d_amd.cli: only the 3D path (--hnc) is built; 1D solute RISM is not.",
              file=sys.stderr)
        return 2
    settings = {"L": a.L, "N": a.N, "rho": a.rho, "beta": a.beta, "norm-tol": a.norm_tol,
                "max-iter": a.max_iter, "closure": a.closure, "dielectric": a.dielectric,
                "lambda": a.lam, "snes-solver": a.snes_solver, "precision": a.precision,
                "verbose": int(a.verbose)}
    if a.rmax is not None:
        settings["rmax"] = a.rmax
    if a.nrad is not None:
        settings["nrad"] = a.nrad
    if a.no_renorm:
        settings["no-renorm"] = True
    solute = a.solute or a.solvent
    chi = None
    if not a.chi_upscale:
        chi = R.rism_solvent(a.solvent, settings)["susceptibility"]
    out = R.hnc3d_run_solute(solute, a.solvent, settings, chi=chi)
    print(out["free-energy"])
    return 0 if out["status"] == "OK" else 1


if __name__ == "__main__":
    sys.exit(main())
