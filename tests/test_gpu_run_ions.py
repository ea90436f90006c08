"""GPU parity on net-charged solutes: the 3D rows of test/out/run-ions-out
through the C ABI (test/run-ions.scm:16-30: nine JC08 ions in cSPC/E water,
--L 10 --N 64, rmax 160, nrad 4096, tau renormalised).

Each closure's golden rows carry one constant of the reference's own chi
run (the ion mu moves by ~6e-3 kcal/mol for a 1e-10 change of chi near
k -> 0; see tests/test_oracle_run_ions.py, which pins the oracle so).  On
the GPU:
* KH with the oracle's chi (committed fixture): the nine ions are one
  constant from the golden to 1.5e-4, and Na+ equals the oracle's 3D solve
  to 1e-7;
* PSE3 with the product's own 1D DRISM chi (rism3d_solvent_chi): one
  constant to 1.5e-4 as well.
"""
import numpy as np
import pytest

import ion_rows
from conftest import CHI_CSPCE_KH, load_chi

pytestmark = pytest.mark.gpu

G = ion_rows.golden()["3d"]
IONS = list(G["KH"])


def _settings(closure):
    return {"N": 64, "L": 10.0, "rmax": ion_rows.RMAX, "nrad": ion_rows.NRAD,
            "rho": ion_rows.RHO, "beta": ion_rows.BETA, "closure": closure,
            "norm-tol": 1e-11, "max-iter": 3000, "device": 0}


def _rows(closure, chi):
    from bgy3d_amd import rism3d as R
    d = {}
    for ion in IONS:
        print(f"run-ions 3D {closure} {ion}", flush=True)
        out = R.hnc3d_run_solute(ion, ion_rows.SOLVENT, _settings(closure), chi=chi)
        assert out["status"] == "OK", (ion, out["status"])
        d[ion] = out["free-energy"]
    return d


def test_kh_rows_and_oracle():
    from oracle import closures as C
    from oracle import rism3d as O
    from bgy3d_amd.solutes import find_molecule
    chi = load_chi(CHI_CSPCE_KH)
    mu = _rows("KH", chi)
    d = np.array([mu[i] - G["KH"][i] for i in IONS])
    print("KH offsets", d, flush=True)
    assert np.max(np.abs(d - d.mean())) <= 1.5e-4, d
    assert abs(d.mean()) < 5e-2
    na = "Na+, JC08, SPC/E"
    pd = O.Problem.from_cli(10.0, 64, ion_rows.BETA, ion_rows.RHO, closure=C.KH,
                            rmax=ion_rows.RMAX, nrad=ion_rows.NRAD)
    ref = O.solute_solve(pd, find_molecule(ion_rows.SOLVENT), find_molecule(na), chi, atol=1e-11)
    print(f"Na+ KH gpu {mu[na]:.10f} oracle {ref['mu']:.10f}", flush=True)
    assert abs(mu[na] - ref["mu"]) < 1e-7


def test_pse3_rows_with_product_chi():
    from bgy3d_amd import rism3d as R
    chi = R.rism_solvent(ion_rows.SOLVENT, {"rmax": ion_rows.RMAX, "nrad": ion_rows.NRAD,
                                            "rho": ion_rows.RHO, "beta": ion_rows.BETA,
                                            "dielectric": 78.4, "closure": "PSE3"})["susceptibility"]
    mu = _rows("PSE3", chi)
    d = np.array([mu[i] - G["PSE3"][i] for i in IONS])
    print("PSE3 offsets", d, flush=True)
    assert np.max(np.abs(d - d.mean())) <= 1.5e-4, d
    assert abs(d.mean()) < 5e-2
