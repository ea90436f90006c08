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

With --solvent-3d (test/Makefile HNC-FLAGS) the solvent's susceptibility
comes from the 3D solvent-solvent solve instead (hnc3d_solvent_solve): run
without --solute it solves the solvent, prints its mu and writes
g%d%d.bin and x%d%d-fft.bin (PETSc binary Vecs) to the current directory;
run with --solute it reads x%d%d-fft.bin back, solves the solute on that
chi, writes g%d.bin and prints the lines print_info and print_chempot
print (bgy3d-potential.c:341-443, hnc3d.c:669-687), the ones
test/summary.sh extracts.
"""
import os
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
    ap.add_argument("--coarse-tol", type=float, default=0.0,
                    help="solve on the half grid to this residual first and start from "
                         "its solution (no counterpart in the reference; 0: off)")
    ap.add_argument("--chi-upscale", action="store_true")
    ap.add_argument("--solvent-3d", action="store_true")
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
                "verbose": int(a.verbose), "coarse-tol": a.coarse_tol}
    if a.rmax is not None:
        settings["rmax"] = a.rmax
    if a.nrad is not None:
        settings["nrad"] = a.nrad
    if a.no_renorm:
        settings["no-renorm"] = True
    if a.solvent_3d:
        return solvent_3d(a, settings)
    solute = a.solute or a.solvent
    chi = None
    if not a.chi_upscale:
        chi = R.rism_solvent(a.solvent, settings)["susceptibility"]
    out = R.hnc3d_run_solute(solute, a.solvent, settings, chi=chi)
    print(out["free-energy"])
    return 0 if out["status"] == "OK" else 1


def print_chempot(out, closure):
    """The three mu lines of print_chempot, '*' on the native closure."""
    for name, key in (("HNC", "free-energy-hnc"), ("KH", "free-energy-kh"),
                      ("GF", "free-energy-gf")):
        star = "*" if name == closure.upper() else ""
        print(f" # mu = {out[key]:f} kcal ({name}){star}")


def print_info(inf):
    """print_info's lines (bgy3d-potential.c:341-443)."""
    def vec(d):
        return f"|{d[0]: f}, {d[1]: f}, {d[2]: f}| = {sum(x * x for x in d) ** 0.5:f}"
    print(f"|<x|ρ_n>| = {vec(inf['d_n'])} (dipole moment of solute cores)")
    print(f"|<x|ρ_u>| = {vec(inf['d_u'])} (dipole moment of solute charge density)")
    print(f"q_u = {inf['q_u']:f} (charge of solute charge density)")
    print(f"|<x|ρ_v>| = {vec(inf['d_v'])} (dipole moment of solvent medium)")
    print(f"q_v = {inf['q_v']:f} (charge of solvent medium)")
    print(f"<U_v|ρ_N> = {inf['U_v.rho_N']:f} (solvent electrostatic field with solute point nuclei)")
    print(f"<U_v|ρ_u> = {inf['U_v.rho_u']:f} "
          "(solvent electrostatic field with diffuse charge density of solute)")
    print(f"<ρ_v|U_u> = {inf['rho_v.U_u']:f} "
          "(solvent charge density with long-range electrostatic field of solute)")


def solvent_3d(a, settings):
    from . import petsc_io
    from .solutes import find_molecule
    m = len(find_molecule(a.solvent))
    if a.solute is None:
        out = R.hnc3d_run_solvent(a.solvent, settings, save=os.getcwd())
        print_chempot(out, a.closure)
        return 0 if out["status"] == "OK" else 1
    N = a.N
    pairs = [(i, j) for i in range(m) for j in range(i + 1)]
    chi = [petsc_io.read_vec(f"x{i}{j}-fft.bin", (N, N, N // 2 + 1, 2))[..., 0] for i, j in pairs]
    out = R.hnc3d_run_solute(a.solute, a.solvent, settings, chi3d=chi, info=True)
    for i in range(m):
        petsc_io.write_vec(f"g{i}.bin", out["GUV"][i])
    if "info" in out:
        print_info(out["info"])
    print_chempot(out, a.closure)
    return 0 if out["status"] == "OK" else 1


if __name__ == "__main__":
    sys.exit(main())
