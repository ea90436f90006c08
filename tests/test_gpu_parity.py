"""GPU parity: the gfx950 kernels through the C ABI against the oracle.

Every case prints its name first so a failing GPU run names it.  Inputs are
seeded; chi comes from the committed fixture (oracle 1D DRISM, README
settings).  Tolerances: FP64 residuals 1e-9 absolute (FFT rounding on
|t| ~ 10); converged mu 1e-8 kcal/mol and g 1e-8 max-abs at 1e-12
residual; MIXED (FP32 bulk) mu within 0.01 kcal/mol of FP64.
"""
import numpy as np
import pytest

from conftest import README
from oracle import closures as C
from oracle import rism3d as O

pytestmark = pytest.mark.gpu

W = "water, PR-SPC/E"


def _R():
    from bgy3d_amd import rism3d as R
    return R


def _sites(name):
    from bgy3d_amd.solutes import find_molecule
    return find_molecule(name)


def _settings(N, L=10.0, closure="HNC", **kw):
    s = dict(README, N=N, L=L, rmax=40.0, nrad=1536, closure=closure, **kw)
    return s


def _oracle_problem(s):
    N = s["N"]
    pd = O.Problem.from_cli(s["L"], N, s["beta"], s["rho"], closure=C.NAMES[s["closure"]],
                            rmax=s["rmax"], nrad=s["nrad"])
    for i, k in enumerate(("nx", "ny", "nz")):
        if k in s:
            n = list(pd.N); n[i] = s[k]; pd.N = tuple(n)
    for i, k in enumerate(("sx", "sy", "sz")):
        if k in s:
            l = list(pd.L); l[i] = s[k]; pd.L = tuple(l)
    pd.h = tuple(l / n for l, n in zip(pd.L, pd.N))
    return pd


def _smooth_t(pd, m, seed):
    rng = np.random.default_rng(seed)
    x, y, z = O.r_axes(pd)
    r2 = z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2
    t = np.stack([rng.uniform(-1, 1) * np.exp(-r2 / rng.uniform(4, 16)) for _ in range(m)])
    return t + 0.01 * rng.standard_normal(t.shape)


RESID_CASES = [
    # name, N (or dict), closure, renorm, solute
    ("n32-hnc-renorm", 32, "HNC", True, W),
    ("n48-kh-renorm-radix3", 48, "KH", True, W),
    ("n64-pse3-norenorm", 64, "PSE3", False, W),
    ("n32-py-renorm", 32, "PY", True, W),
    ("n32-na-kh-charged", 32, "KH", True, "Na+, JC08, SPC/E"),
    ("noncubic-32x48x64", {"N": 32, "ny": 48, "nz": 64, "sy": 24.0, "sz": 28.0}, "HNC", True, W),
]


@pytest.mark.parametrize("name,N,closure,renorm,solute", RESID_CASES, ids=[c[0] for c in RESID_CASES])
def test_residual_fp64(name, N, closure, renorm, solute, chi_w):
    print(f"case residual-fp64 {name}", flush=True)
    R = _R()
    if isinstance(N, dict):
        s = _settings(N["N"], closure=closure, **{k: v for k, v in N.items() if k != "N"})
    else:
        s = _settings(N, closure=closure)
    if not renorm:
        s["no-renorm"] = True
    pd = _oracle_problem(s)
    ref = O.Solute3D(pd, _sites(W), _sites(solute), chi_w, renorm=renorm)
    t = _smooth_t(pd, 3, seed=sum(name.encode()))
    r_ref = ref.iterate(t)
    r_gpu = R.residual(solute, W, s, chi_w, t)
    err = np.max(np.abs(r_gpu - r_ref))
    scale = np.max(np.abs(r_ref))
    print(f"  max|dF| = {err:.3e}  max|F| = {scale:.3e}", flush=True)
    assert err < 1e-9 * max(1.0, scale)


def test_residual_mixed(chi_w):
    print("case residual-mixed n64-hnc", flush=True)
    R = _R()
    s = _settings(64, precision="mixed")
    pd = _oracle_problem(s)
    ref = O.Solute3D(pd, _sites(W), _sites(W), chi_w)
    t = _smooth_t(pd, 3, seed=7)
    r_ref = ref.iterate(t)
    r_gpu = R.residual(W, W, s, chi_w, t)
    err = np.max(np.abs(r_gpu - r_ref))
    print(f"  max|dF| = {err:.3e}", flush=True)
    assert err < 2e-5 * max(1.0, np.max(np.abs(r_ref)))


@pytest.mark.parametrize("N", [64, 96, 128])
def test_residual_mixed_planes(chi_w, N):
    """The FP32 residual on xy planes (three launches, the default for square
    planes of 64-128 points) against the five-launch chain
    (RISM3D_PLANES=0) and the oracle, charged 18-site solute."""
    import os
    print(f"case residual-mixed planes n{N}-kh charged", flush=True)
    R = _R()
    from bgy3d_amd import synthetic as S
    sol = S.solute(5)
    s = _settings(N, L=10.0, closure="KH", precision="mixed")
    pd = _oracle_problem(s)
    ref = O.Solute3D(pd, _sites(W), sol, chi_w)
    t = _smooth_t(pd, 3, seed=11)
    r_ref = ref.iterate(t)
    old = os.environ.get("RISM3D_PLANES")
    try:
        os.environ["RISM3D_PLANES"] = "1"
        r_pl = R.residual(sol, W, s, chi_w, t)
        os.environ["RISM3D_PLANES"] = "0"
        r_rows = R.residual(sol, W, s, chi_w, t)
    finally:
        if old is None:
            os.environ.pop("RISM3D_PLANES", None)
        else:
            os.environ["RISM3D_PLANES"] = old
    scale = max(1.0, np.max(np.abs(r_ref)))
    e_pl, e_rows = np.max(np.abs(r_pl - r_ref)), np.max(np.abs(r_rows - r_ref))
    e_pr = np.max(np.abs(r_pl - r_rows))
    print(f"  max|F| {scale:.3e} planes-oracle {e_pl:.3e} rows-oracle {e_rows:.3e} "
          f"planes-rows {e_pr:.3e}", flush=True)
    assert e_pl < 2e-5 * scale and e_rows < 2e-5 * scale
    assert e_pr < 2e-5 * scale


SOLVE_CASES = [
    ("n32-hnc", 32, "HNC", True, W),
    ("n48-kh", 48, "KH", True, W),
    ("n32-pse3", 32, "PSE3", True, W),
    ("n32-hnc-norenorm", 32, "HNC", False, W),
    ("n32-na-kh", 32, "KH", True, "Na+, JC08, SPC/E"),
    ("n32-cl-pse2", 32, "PSE2", True, "Cl-, JC08, SPC/E"),
]


@pytest.mark.parametrize("name,N,closure,renorm,solute", SOLVE_CASES, ids=[c[0] for c in SOLVE_CASES])
def test_solve_fp64(name, N, closure, renorm, solute, chi_w):
    print(f"case solve-fp64 {name}", flush=True)
    R = _R()
    s = _settings(N, closure=closure, **{"norm-tol": 1e-12, "max-iter": 2000})
    if not renorm:
        s["no-renorm"] = True
    pd = _oracle_problem(s)
    ref = O.solute_solve(pd, _sites(W), _sites(solute), chi_w, renorm=renorm, atol=1e-12)
    out = R.hnc3d_run_solute(solute, W, s, chi=chi_w)
    dmu = abs(out["free-energy"] - ref["mu"])
    dg = np.max(np.abs(out["GUV"] - ref["g"]))
    print(f"  mu gpu {out['free-energy']:.10f} oracle {ref['mu']:.10f} |dmu| {dmu:.2e} "
          f"max|dg| {dg:.2e} iters {out['iterations']}", flush=True)
    assert out["status"] == "OK"
    assert dmu < 1e-8
    assert dg < 1e-8
    for key, ok in (("free-energy-hnc", "mu_hnc"), ("free-energy-kh", "mu_kh"),
                    ("free-energy-gf", "mu_gf")):
        assert abs(out[key] - ref[ok]) < 1e-8
    assert abs(out["partial-molar-volume-factor"] - ref["pmv_factor"]) < 1e-8


def test_readme_anchor_96(chi_w):
    """README:20-31: mu = -6.349052 (renorm), -6.297170 (--no-renorm);
    oracle values -6.3490516355 / -6.2971696813 on the same chi."""
    print("case readme-96", flush=True)
    R = _R()
    for renorm, readme, oracle_mu in ((True, -6.349052, -6.349051635500067),
                                      (False, -6.297170, -6.297169681334508)):
        s = _settings(96, **{"norm-tol": 1e-12, "max-iter": 2000})
        if not renorm:
            s["no-renorm"] = True
        out = R.hnc3d_run_solute(W, W, s, chi=chi_w)
        mu = out["free-energy"]
        print(f"  renorm={renorm} mu {mu:.10f} README {readme} iters {out['iterations']}", flush=True)
        assert abs(mu - readme) < 1e-6
        assert abs(mu - oracle_mu) < 1e-8


def test_mixed_vs_fp64(chi_w):
    print("case mixed-vs-fp64 n64", flush=True)
    R = _R()
    s = _settings(64, **{"norm-tol": 1e-10, "max-iter": 2000})
    f64 = R.hnc3d_run_solute(W, W, s, chi=chi_w)
    for polish in (0, 3):
        m = R.hnc3d_run_solute(W, W, dict(s, precision="mixed", **{"polish-iters": polish}), chi=chi_w)
        dmu = abs(m["free-energy"] - f64["free-energy"])
        dg = np.max(np.abs(m["GUV"] - f64["GUV"]))
        print(f"  polish {polish}: |dmu| {dmu:.2e} max|dg| {dg:.2e} iters {m['iterations']} "
              f"fp32 {m['iterations-fp32']} residual {m['residual']:.2e} {m['status']}", flush=True)
        assert m["status"] == "OK"
        assert dmu < 0.01


def test_batch_equals_single(chi_w):
    print("case batch-vs-single n32", flush=True)
    R = _R()
    s = _settings(32, closure="KH", **{"norm-tol": 1e-11, "max-iter": 2000})
    solutes = [W, "Na+, JC08, SPC/E", "Cl-, JC08, SPC/E", "OW2"]
    batch = R.batch_solve(solutes, W, s, chi_w)
    for name, b in zip(solutes, batch):
        one = R.hnc3d_run_solute(name, W, s, chi=chi_w)
        print(f"  {name}: batch {b['free-energy']:.10f} single {one['free-energy']:.10f}", flush=True)
        assert b["status"] == "OK"
        assert abs(b["free-energy"] - one["free-energy"]) < 1e-9


def test_restart_token(chi_w):
    print("case restart n32", flush=True)
    R = _R()
    s = _settings(32, **{"norm-tol": 1e-11, "max-iter": 2000})
    a = R.hnc3d_run_solute(W, W, s, chi=chi_w, restart=0)
    tok = a["RESTART"]
    assert tok
    b = R.hnc3d_run_solute(W, W, s, chi=chi_w, restart=tok)
    print(f"  cold iters {a['iterations']} warm iters {b['iterations']}", flush=True)
    assert b["iterations"] < a["iterations"]
    assert abs(a["free-energy"] - b["free-energy"]) < 1e-9
    R.restart_destroy(b["RESTART"])


def test_nonconverged_status(chi_w):
    print("case max-iter-1", flush=True)
    R = _R()
    out = R.hnc3d_run_solute(W, W, _settings(32, **{"max-iter": 1}), chi=chi_w)
    assert out["status"] == "NONCONVERGED"


def test_batch_reports_per_solute_nan(chi_w):
    """HNC diverges for strongly charged solutes (doc/log.txt:96-100): the
    diverging member gets NAN, the others still converge."""
    print("case batch-nan n32", flush=True)
    R = _R()
    s = _settings(32, closure="HNC", **{"norm-tol": 1e-10, "max-iter": 500})
    out = R.batch_solve([W, "Li+, JC08, SPC/E"], W, s, chi_w)
    print(f"  statuses {[o['status'] for o in out]}", flush=True)
    assert out[0]["status"] == "OK"
    assert out[1]["status"] in ("OK", "NAN")


def test_readme_energy_hnc_end_to_end():
    """README:20-28 as the reference runs it: chi is not supplied, so the
    solver builds it with 1D DRISM (eps 78.4) on the upscaled grid
    rmax = 40, nrad = 1536 (hnc3d.c:772-783, 1471-1495), then solves."""
    print("case readme-96-chi-null", flush=True)
    R = _R()
    s = {"N": 96, "L": 10.0, "rho": 0.0333295, "beta": 1.6889, "dielectric": 78.4,
         "norm-tol": 1e-12, "max-iter": 2000}
    out = R.hnc3d_run_solute(W, W, s, chi=None)
    print(f"  mu {out['free-energy']:.10f} README -6.349052", flush=True)
    assert abs(out["free-energy"] - (-6.349052)) < 1e-6
    assert abs(out["free-energy"] - (-6.349051635500067)) < 1e-8


def test_cli_energy_hnc(capsys):
    """The README command through the energy --hnc equivalent."""
    print("case cli-energy-hnc", flush=True)
    from bgy3d_amd import cli
    rc = cli.main(["energy", "--hnc", "--norm-tol", "1e-12", "--dielectric", "78.4",
                   "--rho", "0.0333295", "--beta", "1.6889", "--L", "10", "--N", "96",
                   "--solvent", "water, PR-SPC/E", "--solute", "water, PR-SPC/E",
                   "--chi-upscale"])
    out = capsys.readouterr().out.strip().splitlines()[-1]
    print(f"  printed {out}", flush=True)
    assert rc == 0
    assert abs(float(out) - (-6.349052)) < 1e-6
    # the same README number from the coarse start (96^3 from 48^3)
    rc = cli.main(["energy", "--hnc", "--norm-tol", "1e-12", "--dielectric", "78.4",
                   "--rho", "0.0333295", "--beta", "1.6889", "--L", "10", "--N", "96",
                   "--solvent", "water, PR-SPC/E", "--solute", "water, PR-SPC/E",
                   "--chi-upscale", "--coarse-tol", "1e-3"])
    out = capsys.readouterr().out.strip().splitlines()[-1]
    print(f"  printed {out} (coarse start)", flush=True)
    assert rc == 0
    assert abs(float(out) - (-6.349052)) < 1e-6


def test_forces_scm_anchors_gpu():
    """test/forces.scm:28-30 through the C ABI with chi_k == NULL: the C++
    1D producer makes chi on the upscaled grid, then the 3D solve."""
    print("case forces-scm-96", flush=True)
    R = _R()
    s = {"N": 96, "L": 10.0, "rho": 0.0333295, "beta": 1.6889, "closure": "HNC",
         "norm-tol": 1e-12, "max-iter": 2000}
    for name, ref in (("OW", 6.86940765594207), ("OW2-UA", 5.82278007673305),
                      ("OW2", 14.0074968008389)):
        out = R.hnc3d_run_solute(name, "OW", s, chi=None)
        print(f"  {name}: mu {out['free-energy']:.12f} reference {ref}", flush=True)
        assert abs(out["free-energy"] - ref) < 1e-8


def test_continuation_hnc(chi_w):
    """A solute whose HNC iteration diverges from t = 0 (synthetic seed 4,
    6 sites): the closure ladder KH, PSE-2..4, HNC reaches the oracle's
    root (which the oracle also reaches without a ladder, at MDIIS step
    0.2, tests/test_oracle_continuation.py)."""
    print("case continuation hnc seed4 n6 n32", flush=True)
    from bgy3d_amd import synthetic as S
    R = _R()
    sol = list(S.solute(4, n=6))
    s = _settings(32, L=8.0, **{"norm-tol": 1e-11, "max-iter": 3000})
    out = R.hnc3d_run_solute(sol, W, s, chi=chi_w)
    ref = O.Solute3D(_oracle_problem(s), _sites(W), sol, chi_w)
    t, _ = ref.solve(atol=1e-11)
    obs = ref.observables(t)
    dmu = abs(out["free-energy"] - obs["mu"])
    dg = np.max(np.abs(out["GUV"] - obs["g"]))
    print(f"  status {out['status']} iters {out['iterations']} |dmu| {dmu:.2e} max|dg| {dg:.2e}",
          flush=True)
    assert out["status"] == "OK"
    assert dmu < 1e-8 and dg < 1e-8
    m = R.hnc3d_run_solute(sol, W, dict(s, precision="mixed"), chi=chi_w)
    dmu = abs(m["free-energy"] - obs["mu"])
    print(f"  mixed: status {m['status']} iters {m['iterations']} |dmu| {dmu:.2e}", flush=True)
    assert m["status"] == "OK" and dmu < 0.01


@pytest.mark.parametrize("dev", ["0", "1"])
def test_graph_replay_matches_eager(chi_w, dev):
    """The residual chain replayed as HIP graphs (default) against the same
    kernels launched one by one (RISM3D_GRAPHS=0): a mixed-precision batch
    (FP32 and FP64 graphs, hundreds of replays with uploads, read-backs and
    rebases between them) gives bit-identical mu and iteration counts, with
    the MDIIS step on the host (default) or on the device (the FP32 graphs
    then end in mdiis_step)."""
    print(f"case graphs vs eager, device MDIIS {dev}", flush=True)
    import os
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sols = [S.solute(s) for s in range(4)]
    st = dict(README, N=32, L=8.0, rmax=40.0, nrad=1536, closure="KH",
              **{"norm-tol": 1e-10, "precision": "mixed"})
    keys = ("RISM3D_GRAPHS", "RISM3D_DEVICE_MDIIS")
    old = {k: os.environ.get(k) for k in keys}
    try:
        os.environ["RISM3D_DEVICE_MDIIS"] = dev
        os.environ["RISM3D_GRAPHS"] = "1"
        a = R.batch_solve(sols, "water, PR-SPC/E", st, chi_w)
        os.environ["RISM3D_GRAPHS"] = "0"
        b = R.batch_solve(sols, "water, PR-SPC/E", st, chi_w)
    finally:
        for k in keys:
            if old[k] is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = old[k]
    for x, y in zip(a, b):
        print(f"  {x['status']} {x['free-energy']!r} {y['free-energy']!r} iters {x['iterations']} {y['iterations']}",
              flush=True)
        assert x["status"] == "OK" and y["status"] == "OK"
        assert x["free-energy"] == y["free-energy"]
        assert x["iterations"] == y["iterations"]


@pytest.mark.gpu
@pytest.mark.parametrize("wait", ["0", "3"])
def test_device_mdiis_matches_host(chi_w, wait):
    """The FP32 phase's MDIIS step on the device (mdiis_step,
    RISM3D_DEVICE_MDIIS=1) against the same step on the host (the default):
    the device solves the DIIS system with the host's operations in the
    host's order,
    so a mixed batch (hundreds of passes, rebases, restarts of the history)
    gives bit-identical mu, g and iteration counts, with rebases one by one
    (RISM3D_REBASE_WAIT=0) or grouped (3: solutes ready for a rebase wait up
    to 3 passes for one another, iterating on)."""
    print(f"case device vs host MDIIS, rebase wait {wait}", flush=True)
    import os
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sols = [S.solute(s) for s in range(6)]
    st = dict(README, N=32, L=8.0, rmax=40.0, nrad=1536, closure="KH",
              **{"norm-tol": 1e-10, "precision": "mixed"})
    keys = ("RISM3D_DEVICE_MDIIS", "RISM3D_REBASE_WAIT")
    old = {k: os.environ.get(k) for k in keys}
    try:
        os.environ["RISM3D_REBASE_WAIT"] = wait
        os.environ["RISM3D_DEVICE_MDIIS"] = "1"
        a = R.batch_solve(sols, "water, PR-SPC/E", st, chi_w, want_g=True)
        os.environ["RISM3D_DEVICE_MDIIS"] = "0"
        b = R.batch_solve(sols, "water, PR-SPC/E", st, chi_w, want_g=True)
    finally:
        for k in keys:
            if old[k] is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = old[k]
    for x, y in zip(a, b):
        print(f"  {x['status']} {x['free-energy']!r} {y['free-energy']!r} iters {x['iterations']} {y['iterations']}",
              flush=True)
        assert x["status"] == "OK" and y["status"] == "OK"
        assert x["free-energy"] == y["free-energy"]
        assert x["iterations"] == y["iterations"]
        assert x["iterations-fp32"] == y["iterations-fp32"]
        assert np.array_equal(x["GUV"], y["GUV"])
