"""Solvent mixtures on the GPU (options.site_rho; no counterpart in the
reference): the symmetric table S = D^-1/2 chi D^1/2 with the per-site
scalings in the z pass, per-site densities in mu.

* water as two copies at 0.3 / 0.7 rho is pure water: mu and g equal the
  pure-water solve to 1e-9 (the invariant of tests/test_oracle_mixture.py);
* water + Na+ / Cl- at 0.1 M (5 sites: the 5-site z-pass instantiation)
  against oracle/rism3d.py Solute3D(site_rho=): mu and g to 1e-8;
* the same, mixed precision: |dmu| < 0.01 against FP64;
* the large mixture workload: 600-site solute (net +4) in that 5-site
  solvent, KH + tau, 384^3 in a 96 A box, mixed against FP64.
"""
import time

import numpy as np
import pytest

from oracle import closures as C
from oracle import rism1d as R1
from oracle import rism3d as O

pytestmark = pytest.mark.gpu
BETA, RHO = 1.6889, 0.0333295
NA = ("Na", (0.0, 0.0, 0.0), 2.35, 0.13, 1.0)
CL = ("Cl", (0.0, 0.0, 0.0), 4.40, 0.10, -1.0)
RHO_ION = 0.1 * 6.02214076e-4


def _water():
    from bgy3d_amd.solutes import find_molecule
    return list(find_molecule("water, PR-SPC/E"))


def _settings(N, L, nrad, rmax, **kw):
    return dict({"N": N, "L": L, "beta": BETA, "rho": RHO, "closure": "KH", "rmax": rmax,
                 "nrad": nrad, "norm-tol": 1e-11, "max-iter": 3000}, **kw)


def _saline_chi(nrad, rmax):
    from bgy3d_amd import rism3d as R
    sites = _water() + [NA, CL]
    rho = [RHO] * 3 + [RHO_ION] * 2
    s = {"rmax": rmax, "nrad": nrad, "beta": BETA, "rho": RHO, "closure": "KH", "L": 8, "N": 32}
    return sites, rho, R.rism_solvent_mixture(sites, [0, 0, 0, 1, 2], rho, s)["susceptibility"]


def test_copies_equal_pure_water():
    print("case mixture copies n32", flush=True)
    from bgy3d_amd import rism3d as R
    W = _water()
    s1 = {"rmax": 40.0, "nrad": 512, "beta": BETA, "rho": RHO, "closure": "KH", "L": 8, "N": 32}
    pure = R.rism_solvent(W, dict(s1, dielectric=0.0))["susceptibility"]
    rho = [0.3 * RHO] * 3 + [0.7 * RHO] * 3
    mix = R.rism_solvent_mixture(W + W, [0, 0, 0, 1, 1, 1], rho, s1)["susceptibility"]
    st = _settings(32, 8.0, 512, 40.0)
    a = R.hnc3d_run_solute("water, cSPC/E", W, st, chi=pure)
    b = R.hnc3d_run_solute("water, cSPC/E", W + W, st, chi=mix, site_rho=rho)
    print(f"  mu pure {a['free-energy']:.10f} copies {b['free-energy']:.10f}", flush=True)
    assert a["status"] == "OK" and b["status"] == "OK"
    assert abs(a["free-energy"] - b["free-energy"]) < 1e-9
    assert np.max(np.abs(b["GUV"][:3] - a["GUV"])) < 1e-9
    assert np.max(np.abs(b["GUV"][3:] - a["GUV"])) < 1e-9


def test_saline_vs_oracle():
    print("case mixture saline n32 vs oracle", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd.solutes import find_molecule
    sites, rho, chi = _saline_chi(512, 40.0)
    sol = find_molecule("water, cSPC/E")
    pd = O.Problem.from_cli(8.0, 32, BETA, RHO, closure=C.KH, rmax=40.0, nrad=512)
    ref = O.Solute3D(pd, sites, sol, chi, site_rho=rho)
    t, _ = ref.solve(atol=1e-11)
    obs = ref.observables(t)
    out = R.hnc3d_run_solute(sol, sites, _settings(32, 8.0, 512, 40.0), chi=chi, site_rho=rho)
    print(f"  {out['status']} mu {out['free-energy']:.10f} oracle {obs['mu']:.10f} "
          f"pmv {out['partial-molar-volume-factor']:.8f} {obs['pmv_factor']:.8f}", flush=True)
    assert out["status"] == "OK"
    assert abs(out["free-energy"] - obs["mu"]) < 1e-8
    assert abs(out["partial-molar-volume-factor"] - obs["pmv_factor"]) < 1e-8
    assert np.max(np.abs(out["GUV"] - obs["g"])) < 1e-8
    mixed = R.hnc3d_run_solute(sol, sites, _settings(32, 8.0, 512, 40.0, precision="mixed"),
                               chi=chi, site_rho=rho)
    print(f"  mixed {mixed['status']} |dmu| {abs(mixed['free-energy'] - obs['mu']):.2e}",
          flush=True)
    assert mixed["status"] == "OK"
    assert abs(mixed["free-energy"] - obs["mu"]) < 0.01


def test_large_mixture_384():
    """600-site synthetic solute, net charge +4, in water + 0.1 M NaCl, KH
    with tau renormalisation, 384^3 grid, 96 A box, 0.25 A spacing."""
    print("case mixture 384^3 n=600 Q=+4", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sites, rho, chi = _saline_chi(4096, 160.0)
    sol = S.solute(7, n=600, net=4)
    st = _settings(384, 48.0, 4096, 160.0, **{"norm-tol": 1e-10})
    t0 = time.perf_counter()
    mixed = R.hnc3d_run_solute(sol, sites, dict(st, precision="mixed"), chi=chi, site_rho=rho)
    t_mixed = time.perf_counter() - t0
    t0 = time.perf_counter()
    fp64 = R.hnc3d_run_solute(sol, sites, st, chi=chi, site_rho=rho)
    t_fp64 = time.perf_counter() - t0
    dmu = abs(mixed["free-energy"] - fp64["free-energy"])
    print(f"  mixed {mixed['status']} mu {mixed['free-energy']:.6f} iters {mixed['iterations']} "
          f"({mixed['iterations-fp32']} fp32) {t_mixed:.2f} s; fp64 {fp64['status']} mu "
          f"{fp64['free-energy']:.6f} iters {fp64['iterations']} {t_fp64:.2f} s; |dmu| {dmu:.2e}; "
          f"max|dg| {np.max(np.abs(mixed['GUV'] - fp64['GUV'])):.2e}", flush=True)
    assert mixed["status"] == "OK" and fp64["status"] == "OK"
    assert dmu < 0.01
