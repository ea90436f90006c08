"""Coarse start (rism3d_options.coarse_tol): the solve on the grid N/2 in
the same box, its t trigonometrically interpolated to N, then the solve on
N.  The root is the cold start's: mu and g agree to what the stopping test
(max|F| <= 1e-10) leaves, on the charged and uncharged paths, with the
tau renormalisation and the Ng form.  A wrong interpolation (scale, Nyquist
split, index map) would still reach the root, only later, so the residual
count on N is checked too: the interpolated start must save residuals.
Where N/2 is not a supported size the option is ignored, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
W = "water, PR-SPC/E"
BASE = {"rho": 0.0333295, "beta": 1.6889, "rmax": 160.0, "nrad": 4096, "max-iter": 3000,
        "norm-tol": 1e-10}


@pytest.fixture(scope="module")
def chi160():
    from conftest import load_chi
    return load_chi("chi_prspce_hnc_eps78.4_rmax160_nrad4096.npy")


CASES = [
    ("KH tau 64^3 charged", dict(N=64, L=12.0, closure="KH"), 3, 12, None, 1.0),
    ("KH tau 128^3 charged", dict(N=128, L=17.25, closure="KH"), 3, 12, None, 0.85),
    ("HNC Ng 96^3 charged", dict(N=96, L=14.0, closure="HNC", **{"no-renorm": True}), 5, 8,
     None, 1.0),
    ("PSE3 tau 96^3 uncharged", dict(N=96, L=14.0, closure="PSE3"), 7, 10, 0, 1.0),
]


@pytest.mark.parametrize("name,over,seed,n,net,frac", CASES, ids=[c[0] for c in CASES])
def test_coarse_start_same_root(chi160, name, over, seed, n, net, frac):
    print(f"case coarse start {name}", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sol = S.solute(seed, n=n, net=net) if net is not None else S.solute(seed, n=n)
    s = dict(BASE, **over)
    cold = R.hnc3d_run_solute(sol, W, s, chi=chi160)
    warm = R.hnc3d_run_solute(sol, W, dict(s, **{"coarse-tol": 1e-3}), chi=chi160)
    dmu = abs(cold["free-energy"] - warm["free-energy"])
    dg = float(np.max(np.abs(cold["GUV"] - warm["GUV"])))
    print(f"  mu {cold['free-energy']:.10f} {warm['free-energy']:.10f} |dmu| {dmu:.1e} "
          f"max|dg| {dg:.1e} residuals {cold['iterations']} -> {warm['iterations']}", flush=True)
    assert cold["status"] == "OK" and warm["status"] == "OK"
    assert dmu < 1e-7 and dg < 1e-7
    assert warm["iterations"] < frac * cold["iterations"]


def test_coarse_start_mixed_batch(chi160):
    """Mixed precision, a batch of 6 on 96^3 (coarse 48^3): every solute
    converges and mu stays within the mixed bar of the cold FP64 solve."""
    print("case coarse start mixed batch 6 x 96^3", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sols = S.batch(range(6))
    box = S.forced_box(sols, N=96)
    s = dict(BASE, N=96, L=box / 2, closure="KH")
    cold = R.batch_solve(sols, W, s, chi160)
    warm = R.batch_solve(sols, W, dict(s, precision="mixed", **{"mixed-tol": 1e-5,
                                                                "polish-iters": 3,
                                                                "coarse-tol": 1e-3}), chi160)
    for a, b in zip(cold, warm):
        print(f"  mu {a['free-energy']:.6f} {b['free-energy']:.6f} residuals "
              f"{a['iterations']} -> {b['iterations']}", flush=True)
        assert a["status"] == "OK" and b["status"] == "OK"
        assert abs(a["free-energy"] - b["free-energy"]) < 0.01


def test_coarse_start_ignored_where_half_grid_unsupported(chi160):
    """N = 48: N/2 = 24 is not a supported size, so the option changes
    nothing."""
    print("case coarse start ignored 48^3", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sol = S.solute(4, n=6)
    s = dict(BASE, N=48, L=10.0, closure="KH")
    a = R.hnc3d_run_solute(sol, W, s, chi=chi160)
    b = R.hnc3d_run_solute(sol, W, dict(s, **{"coarse-tol": 1e-3}), chi=chi160)
    assert a["free-energy"] == b["free-energy"] and a["iterations"] == b["iterations"]
    assert np.array_equal(a["GUV"], b["GUV"])


def test_history_widths_reach_one_root(chi160):
    """MDIIS depths 8, 10 and 12 run the closure and residual kernels with
    8, 10 and 16 history slots unrolled; FP64 solves reach one root, and the
    mixed solve at depth 10 stays within the mixed bar."""
    print("case history widths 8 / 10 / 16 at 64^3", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sol = S.solute(11, n=9)
    s = dict(BASE, N=64, L=12.0, closure="KH")
    runs = {d: R.hnc3d_run_solute(sol, W, dict(s, **{"diis-depth": d}), chi=chi160)
            for d in (8, 10, 12)}
    mixed = R.hnc3d_run_solute(sol, W, dict(s, precision="mixed", **{"diis-depth": 10}),
                               chi=chi160)
    for d, r in runs.items():
        print(f"  depth {d}: mu {r['free-energy']:.10f} residuals {r['iterations']}", flush=True)
        assert r["status"] == "OK"
        assert abs(r["free-energy"] - runs[8]["free-energy"]) < 1e-7
        assert float(np.max(np.abs(r["GUV"] - runs[8]["GUV"]))) < 1e-7
    print(f"  mixed depth 10: mu {mixed['free-energy']:.8f}", flush=True)
    assert mixed["status"] == "OK"
    assert abs(mixed["free-energy"] - runs[8]["free-energy"]) < 0.01


def test_coarse_start_one_site_solvent_forces_anchors():
    """A 1-site solvent runs the all-sites-in-LDS zstar (inverse-only mode
    for the interpolated start): test/forces.scm:28-30 from the coarse start
    (96^3 from 48^3, HNC, OW solvent, chi from the 1D producer) to the same
    1e-8 as the cold start.  These uncharged solutes converge in about 25
    residuals from t = 0, so the start saves little (25 -> 22 for OW, 27 ->
    29 for OW2-UA): only the root is checked."""
    print("case coarse start forces-scm-96 OW solvent", flush=True)
    from bgy3d_amd import rism3d as R
    s = {"N": 96, "L": 10.0, "rho": 0.0333295, "beta": 1.6889, "closure": "HNC",
         "norm-tol": 1e-12, "max-iter": 2000}
    for name, ref in (("OW", 6.86940765594207), ("OW2-UA", 5.82278007673305),
                      ("OW2", 14.0074968008389)):
        cold = R.hnc3d_run_solute(name, "OW", s, chi=None)
        warm = R.hnc3d_run_solute(name, "OW", dict(s, **{"coarse-tol": 1e-3}), chi=None)
        print(f"  {name}: mu {warm['free-energy']:.12f} reference {ref} residuals "
              f"{cold['iterations']} -> {warm['iterations']}", flush=True)
        assert warm["status"] == "OK"
        assert abs(warm["free-energy"] - ref) < 1e-8


def test_interpolated_start_matches_checker(chi160):
    """The start the coarse stage hands to N, point by point: a solve with
    norm-tol 1e9 stops at its first residual and returns that start, which
    must be the trigonometric interpolation (tests/trig_upsample.py, pinned
    on band-limited fields) of the standalone N/2 solve to the same
    coarse_tol."""
    print("case coarse start interpolated t vs checker 32^3 -> 64^3", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    from trig_upsample import upsample
    sol = S.solute(3, n=12)
    s = dict(BASE, N=64, L=12.0, closure="KH")
    coarse = R.hnc3d_run_solute(sol, W, dict(s, N=32, **{"norm-tol": 1e-3}), chi=chi160,
                                want_t=True)
    start = R.hnc3d_run_solute(sol, W, dict(s, **{"norm-tol": 1e9, "coarse-tol": 1e-3}),
                               chi=chi160, want_t=True)
    want = upsample(coarse["t"])
    err = float(np.max(np.abs(start["t"] - want)))
    scale = float(np.max(np.abs(want)))
    print(f"  coarse residuals {coarse['iterations']}, start returned after "
          f"{start['iterations']}; max|t - interp| {err:.2e} of {scale:.2e}", flush=True)
    assert coarse["status"] == "OK" and start["status"] == "OK"
    assert start["iterations"] <= 1
    assert err <= 1e-12 * scale * 100
