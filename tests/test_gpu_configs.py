"""Large single-solute workloads as GPU cases, and the plain FP64 water
case at 96^3 against the oracle.

Precision study: one 120-site synthetic solute, net charge -2,
PSE-3, --L 32 --N 256 (64 A box, h = 0.25 A), PR-SPC/E water.  |dmu| and
max|dg| against the FP64 solve of: FP32 only (FP32 passes around the
starting point until their floor; reported, no bar: the floor is the point
of the study), and the mixed path (defect correction, then 3 FP64 polish
steps; bar |dmu| < 0.01 kcal/mol).  There is no published mu for this case
(parity of the FP64 path to the oracle is covered on smaller grids).

Gradients: one 30-site solute, HNC, --L 20 --N 128; adjoint
d mu / d theta against central finite differences of the GPU mu (h = 1e-4)
to 1e-4 relative, and the time of forward + adjoint against forward alone.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
W = "water, PR-SPC/E"
BASE = {"rho": 0.0333295, "beta": 1.6889, "rmax": 160.0, "nrad": 4096, "max-iter": 3000}


@pytest.fixture(scope="module")
def chi160():
    from conftest import load_chi
    return load_chi("chi_prspce_hnc_eps78.4_rmax160_nrad4096.npy")


def test_precision_study_256_pse3_120_sites(chi160):
    print("case precision study 256^3 PSE3 120 sites Q=-2", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sol = S.solute(2, n=120, net=-2)
    s = dict(BASE, N=256, L=32.0, closure="PSE3", **{"norm-tol": 1e-9})
    t0 = time.time()
    f64 = R.hnc3d_run_solute(sol, W, s, chi=chi160)
    t64 = time.time() - t0
    assert f64["status"] == "OK"
    rows = []
    for name, prec, polish in (("fp32-only", "fp32", 0), ("mixed+3xfp64", "mixed", 3)):
        t0 = time.time()
        m = R.hnc3d_run_solute(sol, W, dict(s, precision=prec, **{"polish-iters": polish,
                                                                   "mixed-tol": 1e-5}),
                               chi=chi160)
        dt = time.time() - t0
        dmu = abs(m["free-energy"] - f64["free-energy"])
        dg = float(np.max(np.abs(m["GUV"] - f64["GUV"])))
        rows.append((name, dmu, dg))
        print(f"  {name}: mu {m['free-energy']:.6f} |dmu| {dmu:.2e} max|dg| {dg:.2e} "
              f"residual {m['residual']:.1e} iters {m['iterations']} "
              f"(fp32 {m['iterations-fp32']}) {dt:.2f} s", flush=True)
        if prec == "mixed":
            assert m["status"] == "OK"
            assert dmu < 0.01
    print(f"  fp64: mu {f64['free-energy']:.6f} iters {f64['iterations']} {t64:.2f} s", flush=True)


def test_adjoint_128_hnc_30_sites(chi160):
    print("case adjoint 128^3 HNC 30 sites", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sol = list(S.solute(3, n=30))
    s = dict(BASE, N=128, L=20.0, closure="HNC", **{"norm-tol": 1e-11})
    t0 = time.time()
    fwd = R.hnc3d_run_solute(sol, W, s, chi=chi160)
    t_fwd = time.time() - t0
    t0 = time.time()
    out = R.hnc3d_run_solute(sol, W, s, chi=chi160, adjoint=True, derivatives=True)
    t_both = time.time() - t0
    g = out["free-energy-dsite"]
    assert np.all(np.isfinite(g))
    print(f"  forward {t_fwd:.2f} s, forward + adjoint {t_both:.2f} s "
          f"(adjoint iters {out['adjoint-iterations']})", flush=True)
    h = 1e-4
    for a, p in ((0, 0), (7, 2), (15, 3), (22, 4), (29, 5)):
        def moved(sign):
            x = [[n, list(xyz), sg, ep, q] for n, xyz, sg, ep, q in sol]
            if p < 3:
                x[a][1][p] += sign * h
            elif p == 3:
                x[a][4] += sign * h
            elif p == 4:
                x[a][2] += sign * h
            else:
                x[a][3] += sign * h
            return [(n, tuple(xyz), sg, ep, q) for n, xyz, sg, ep, q in x]
        fd = (R.hnc3d_run_solute(moved(+1), W, s, chi=chi160)["free-energy"]
              - R.hnc3d_run_solute(moved(-1), W, s, chi=chi160)["free-energy"]) / (2 * h)
        print(f"  site {a} param {p}: adjoint {g[a, p]:.8f} fd {fd:.8f}", flush=True)
        assert abs(g[a, p] - fd) <= 1e-4 * max(abs(fd), 1e-2)


def test_adjoint_128_hnc_30_sites_mixed(chi160):
    """The mixed-precision leg of the same workload: mixed forward solve
    and the adjoint by FP32 defect correction, against central differences
    of the FP64 mu (h = 1e-4) to 1e-4 relative; forward + backward timed
    against forward alone."""
    print("case adjoint 128^3 HNC 30 sites mixed", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sol = list(S.solute(3, n=30))
    s64 = dict(BASE, N=128, L=20.0, closure="HNC", **{"norm-tol": 1e-11})
    sm = dict(s64, precision="mixed", **{"norm-tol": 1e-10})
    R.hnc3d_run_solute(sol, W, sm, chi=chi160)  # warm
    t0 = time.time()
    fwd = R.hnc3d_run_solute(sol, W, sm, chi=chi160)
    t_fwd = time.time() - t0
    t0 = time.time()
    out = R.hnc3d_run_solute(sol, W, sm, chi=chi160, adjoint=True)
    t_both = time.time() - t0
    assert fwd["status"] == "OK" and out["status"] == "OK"
    g = out["free-energy-dsite"]
    assert np.all(np.isfinite(g))
    print(f"  mixed: forward {t_fwd:.2f} s, forward + adjoint {t_both:.2f} s "
          f"(adjoint iters {out['adjoint-iterations']})", flush=True)
    h = 1e-4
    for a, p in ((0, 0), (7, 2), (15, 3), (22, 4), (29, 5)):
        def moved(sign):
            x = [[n, list(xyz), sg, ep, q] for n, xyz, sg, ep, q in sol]
            if p < 3:
                x[a][1][p] += sign * h
            elif p == 3:
                x[a][4] += sign * h
            elif p == 4:
                x[a][2] += sign * h
            else:
                x[a][3] += sign * h
            return [(n, tuple(xyz), sg, ep, q) for n, xyz, sg, ep, q in x]
        fd = (R.hnc3d_run_solute(moved(+1), W, s64, chi=chi160)["free-energy"]
              - R.hnc3d_run_solute(moved(-1), W, s64, chi=chi160)["free-energy"]) / (2 * h)
        print(f"  site {a} param {p}: mixed adjoint {g[a, p]:.8f} fd {fd:.8f}", flush=True)
        assert abs(g[a, p] - fd) <= 1e-4 * max(abs(fd), 1e-2)


def test_config0_water_96_fp64(chi160):
    """The plain FP64 workload as written: PR-SPC/E water in PR-SPC/E water,
    HNC, chi from the 1D run at rmax 160 A / nrad 4096, --L 10 --N 96,
    against the oracle's solve of the same problem to 1e-13
    (tests/golden/config0_water_96.json, written by
    tests/golden/make_config0.py): mu to 1e-9 kcal/mol, and per solvent
    site the sum of g over the grid, max g and g at three grid points to
    1e-9 relative.  The solve runs to max|F| <= 5e-14: the FP64 max-norm
    residual of this grid levels off at about 3e-14 (3000 MDIIS steps
    reach 2.8e-14, not 1e-14); the workload's 1e-14 is the reference's
    relative 2-norm test of its default Newton solver (bgy3d-snes.c)."""
    print("case config0 water/water HNC 96^3 rmax160 nrad4096 norm-tol 5e-14", flush=True)
    import json
    import os
    from bgy3d_amd import rism3d as R
    with open(os.path.join(os.path.dirname(__file__), "golden", "config0_water_96.json")) as f:
        gold = json.load(f)
    s = dict(BASE, N=96, L=10.0, closure="HNC", **{"norm-tol": 5e-14})
    out = R.hnc3d_run_solute(W, W, s, chi=chi160)
    print(f"  mu {out['free-energy']:.12f} oracle {gold['mu']:.12f} residual "
          f"{out['residual']:.2e} iters {out['iterations']} {out['status']}", flush=True)
    assert out["status"] == "OK"
    assert abs(out["free-energy"] - gold["mu"]) < 1e-9
    g = out["GUV"]
    for i, site in enumerate(gold["sites"]):
        assert abs(g[i].sum() - site["sum"]) <= 1e-9 * abs(site["sum"])
        assert abs(g[i].max() - site["max"]) <= 1e-9 * abs(site["max"])
        for p, v in zip(gold["points"], site["at"]):
            assert abs(g[i][tuple(p)] - v) <= 1e-9 * max(1.0, abs(v))
