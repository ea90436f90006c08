"""Size-independent properties at the benchmark's full size (128^3, the
batched workload's grid and box), where the oracle would take minutes:

* exchanging two coordinate axes of the solute exchanges the axes of g
  and leaves mu (cubic grid);
* the Jacobian is linear: J(a v + b w) = a J v + b J w;
* the residual of the converged t is at the tolerance when re-evaluated
  from the returned t (the solver's stopping test, through rism3d_residual).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
W = "water, PR-SPC/E"


def _settings(L, **kw):
    return dict({"N": 128, "L": L / 2, "rho": 0.0333295, "beta": 1.6889, "rmax": 160.0,
                 "nrad": 4096, "closure": "KH", "norm-tol": 1e-10, "max-iter": 3000}, **kw)


@pytest.fixture(scope="module")
def chi160():
    from conftest import load_chi
    return load_chi("chi_prspce_hnc_eps78.4_rmax160_nrad4096.npy")


@pytest.fixture(scope="module")
def case():
    from bgy3d_amd import synthetic as S
    sols = S.batch(range(64))
    return S.solute(0), S.forced_box(sols, N=128)


def test_axis_permutation(chi160, case):
    """Exchanging the solute's x and y (and y and z) coordinates on the
    cubic grid exchanges the axes of g and leaves mu: x is the r2c axis, y
    and z the c2c ones, so an indexing slip in one axis's pass breaks it.
    (A translation by whole grid cells is not a symmetry here: the
    reference forms the short-range field by plain distances in the box,
    without periodic images, and so does the oracle.)"""
    print("case axis permutation 128^3", flush=True)
    from bgy3d_amd import rism3d as R
    sol, L = case
    a = R.hnc3d_run_solute(sol, W, _settings(L), chi=chi160)
    assert a["status"] == "OK"
    for perm, axes in (((1, 0, 2), (2, 3)), ((0, 2, 1), (1, 2))):
        moved = [(n, tuple(x[d] for d in perm), s, e, q) for n, x, s, e, q in sol]
        b = R.hnc3d_run_solute(moved, W, _settings(L), chi=chi160)
        ga = np.swapaxes(a["GUV"], *axes)
        dg = float(np.max(np.abs(ga - b["GUV"])))
        print(f"  perm {perm}: mu {a['free-energy']:.10f} {b['free-energy']:.10f} max|dg| {dg:.2e}",
              flush=True)
        assert b["status"] == "OK"
        assert abs(a["free-energy"] - b["free-energy"]) < 1e-8
        assert dg < 1e-8


def test_jacobian_linear(chi160, case):
    print("case jacobian linearity 128^3", flush=True)
    from bgy3d_amd import rism3d as R
    sol, L = case
    rng = np.random.default_rng(3)
    shape = (3, 128, 128, 128)
    t = 0.1 * rng.standard_normal(shape)
    v = rng.standard_normal(shape)
    w = rng.standard_normal(shape)
    st = _settings(L)
    jv = R.jacobian(sol, W, st, chi160, t, v)
    jw = R.jacobian(sol, W, st, chi160, t, w)
    jc = R.jacobian(sol, W, st, chi160, t, 2.0 * v - 0.5 * w)
    err = float(np.max(np.abs(jc - (2.0 * jv - 0.5 * jw))))
    scale = float(np.max(np.abs(jc)))
    print(f"  max|J(2v - w/2) - (2Jv - Jw/2)| {err:.2e} of {scale:.2e}", flush=True)
    assert err <= 1e-12 * scale * 100


def test_returned_t_is_converged(chi160, case):
    print("case residual of returned t 128^3", flush=True)
    from bgy3d_amd import rism3d as R
    sol, L = case
    st = _settings(L)
    out = R.hnc3d_run_solute(sol, W, st, chi=chi160, want_t=True)
    r = R.residual(sol, W, st, chi160, out["t"])
    print(f"  solver residual {out['residual']:.2e}, re-evaluated {np.max(np.abs(r)):.2e}",
          flush=True)
    assert out["status"] == "OK"
    assert np.max(np.abs(r)) <= 1e-10
    assert abs(np.max(np.abs(r)) - out["residual"]) <= 1e-12
