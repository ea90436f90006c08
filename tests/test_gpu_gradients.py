"""GPU gradients against the oracle.

* TPT gradient (the reference's --derivatives, hnc3d.c:1666-1742) vs
  oracle/gradients.py tpt_gradient: 1e-7 relative to the largest component
  (the grid sum cancels terms up to ~1e4 to a net ~0.1; GPU and numpy sum
  in different orders).
* Exact adjoint d mu / d(x, y, z, q, sigma, epsilon) vs the oracle's
  adjoint (itself checked against finite differences of mu on the CPU,
  tests/test_oracle_gradients.py) to 1e-6 relative, and one direct central
  finite difference of the GPU mu (h = 1e-4) to 1e-4 relative.
* Linear-response gradient (the reference's --response, hnc3d.c:1969-2134),
  computed on the GPU from one adjoint solve, vs the oracle's literal
  restatement with 3 n linear solves: 1e-6 relative, on both the tau and
  the Ng path, and for an uncharged solute vs the adjoint's x, y, z.
"""
import numpy as np
import pytest

from conftest import README
from oracle import closures as C
from oracle import gradients as Gr
from oracle import rism3d as O

pytestmark = pytest.mark.gpu
W = "water, PR-SPC/E"


def _solute():
    from bgy3d_amd import synthetic as S
    return list(S.solute(5, n=4))


def _settings(closure):
    return dict(README, N=32, L=8.0, rmax=40.0, nrad=1536, closure=closure,
                **{"norm-tol": 1e-12, "max-iter": 2000})


@pytest.mark.parametrize("closure", ["HNC", "KH"])
def test_gradients_vs_oracle(closure, chi_w):
    print(f"case gradients {closure}", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd.solutes import find_molecule
    sol = _solute()
    s = _settings(closure)
    out = R.hnc3d_run_solute(sol, W, s, chi=chi_w, derivatives=True, adjoint=True)
    pd = O.Problem.from_cli(8.0, 32, README["beta"], README["rho"], closure=C.NAMES[closure],
                            rmax=40.0, nrad=1536)
    ref = O.Solute3D(pd, find_molecule(W), sol, chi_w)
    t, _ = ref.solve(atol=1e-12)
    obs = ref.observables(t)
    tpt = Gr.tpt_gradient(pd, find_molecule(W), sol, obs["g"])
    adj, _ = Gr.adjoint_gradient(ref, t)
    g_gpu = out["free-energy-gradient"]
    d_gpu = out["free-energy-dsite"]
    e_tpt = np.max(np.abs(g_gpu - tpt)) / np.max(np.abs(tpt))
    # relative to the largest value of each parameter kind (x, y, z share one)
    scale = np.array([np.max(np.abs(adj[:, :3]))] * 3 + [np.max(np.abs(adj[:, p])) for p in (3, 4, 5)])
    e_adj = np.max(np.abs(d_gpu - adj) / scale[None, :])
    print(f"  TPT rel {e_tpt:.2e}  adjoint rel {e_adj:.2e}  adj iters {out['adjoint-iterations']}",
          flush=True)
    assert e_tpt < 1e-7
    assert e_adj < 1e-6


@pytest.mark.parametrize("closure", ["HNC", "KH"])
def test_adjoint_ng_path_vs_oracle(closure, chi_w):
    """The adjoint on the Ng path (--no-renorm): its long-range weight
    (q_i + sum_j (chi - 1)_ij q_j) 4 pi e^{-k^2/4G^2}/k^2 against the
    oracle's, itself checked against finite differences on the CPU."""
    print(f"case adjoint ng {closure}", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd.solutes import find_molecule
    sol = _solute()
    s = dict(_settings(closure), **{"no-renorm": True})
    out = R.hnc3d_run_solute(sol, W, s, chi=chi_w, adjoint=True)
    pd = O.Problem.from_cli(8.0, 32, README["beta"], README["rho"], closure=C.NAMES[closure],
                            rmax=40.0, nrad=1536)
    ref = O.Solute3D(pd, find_molecule(W), sol, chi_w, renorm=False)
    t, _ = ref.solve(atol=1e-12)
    adj, _ = Gr.adjoint_gradient(ref, t)
    d_gpu = out["free-energy-dsite"]
    scale = np.array([np.max(np.abs(adj[:, :3]))] * 3 + [np.max(np.abs(adj[:, p])) for p in (3, 4, 5)])
    e_adj = np.max(np.abs(d_gpu - adj) / scale[None, :])
    print(f"  adjoint rel {e_adj:.2e}  adj iters {out['adjoint-iterations']}", flush=True)
    assert e_adj < 1e-6


@pytest.mark.parametrize("renorm", [True, False])
def test_mixed_adjoint_vs_fp64(renorm, chi_w):
    """After a MIXED solve the adjoint's linear solve runs by FP32 defect
    correction: the same gradient as the FP64 path to the accuracy of the
    mixed forward solution (1e-4 relative per parameter kind)."""
    print(f"case mixed adjoint renorm={renorm}", flush=True)
    from bgy3d_amd import rism3d as R
    sol = _solute()
    s = dict(_settings("KH"), **{"no-renorm": not renorm})
    a = R.hnc3d_run_solute(sol, W, s, chi=chi_w, adjoint=True)["free-energy-dsite"]
    out = R.hnc3d_run_solute(sol, W, dict(s, precision="mixed"), chi=chi_w, adjoint=True)
    b = out["free-energy-dsite"]
    scale = np.array([np.max(np.abs(a[:, :3]))] * 3 + [np.max(np.abs(a[:, p])) for p in (3, 4, 5)])
    e = np.max(np.abs(a - b) / scale[None, :])
    print(f"  mixed vs fp64 rel {e:.2e}  adj iters {out['adjoint-iterations']}", flush=True)
    assert e < 1e-4


def test_adjoint_vs_finite_difference(chi_w):
    print("case adjoint-fd", flush=True)
    from bgy3d_amd import rism3d as R
    sol = _solute()
    s = _settings("HNC")
    out = R.hnc3d_run_solute(sol, W, s, chi=chi_w, adjoint=True)
    h = 1e-4
    for a, p in ((0, 0), (1, 3), (2, 4), (3, 5)):
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
        fd = (R.hnc3d_run_solute(moved(+1), W, s, chi=chi_w)["free-energy"]
              - R.hnc3d_run_solute(moved(-1), W, s, chi=chi_w)["free-energy"]) / (2 * h)
        adj = out["free-energy-dsite"][a, p]
        print(f"  site {a} param {p}: adjoint {adj:.8f} fd {fd:.8f}", flush=True)
        assert abs(adj - fd) <= 1e-4 * max(abs(fd), 1e-2)


@pytest.mark.parametrize("closure,renorm", [("KH", True), ("HNC", False)])
def test_response_vs_oracle(closure, renorm, chi_w):
    print(f"case response {closure} renorm={renorm}", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd.solutes import find_molecule
    sol = find_molecule("water, cSPC/E")
    s = _settings(closure)
    if not renorm:
        s["no-renorm"] = True
    out = R.hnc3d_run_solute(sol, W, dict(s, response=True), chi=chi_w)
    pd = O.Problem.from_cli(8.0, 32, README["beta"], README["rho"], closure=C.NAMES[closure],
                            rmax=40.0, nrad=1536)
    ref = O.Solute3D(pd, find_molecule(W), sol, chi_w, renorm=renorm)
    t, _ = ref.solve(atol=1e-12)
    lit = Gr.response_gradient(ref, t)
    e = np.max(np.abs(out["free-energy-response"] - lit)) / np.max(np.abs(lit))
    print(f"  response rel {e:.2e}  adj iters {out['adjoint-iterations']}", flush=True)
    assert e < 1e-6


def test_response_uncharged_equals_adjoint(chi_w):
    print("case response uncharged", flush=True)
    from bgy3d_amd import rism3d as R
    sol = [("A", (0.0, 0.0, 0.9), 3.16, 0.155, 0.0), ("B", (0.3, 0.1, -0.9), 3.0, 0.1, 0.0)]
    out = R.hnc3d_run_solute(sol, W, _settings("HNC"), chi=chi_w, response=True, adjoint=True,
                             derivatives=True)
    r, d = out["free-energy-response"], out["free-energy-dsite"][:, :3]
    e = np.max(np.abs(r - d)) / np.max(np.abs(d))
    print(f"  response vs adjoint rel {e:.2e}", flush=True)
    assert e < 1e-9
    # uncharged, the mu functional is stationary, and the TPT gradient
    # rho h^3 sum_i <g_i | dv_i> is exact as well
    e_tpt = np.max(np.abs(out["free-energy-gradient"] - d)) / np.max(np.abs(d))
    print(f"  TPT vs adjoint rel {e_tpt:.2e}", flush=True)
    assert e_tpt < 1e-9


def test_autograd_solute_mu(chi_w):
    """torch.autograd through bgy3d_amd.autograd.solute_mu: backward equals
    the adjoint columns and central differences of the forward (h = 1e-4)
    to 1e-4 relative."""
    print("case autograd solute_mu", flush=True)
    import torch
    from bgy3d_amd import autograd as A
    names, p0 = A.params_from_sites(_solute())
    s = _settings("HNC")
    p = torch.tensor(p0, requires_grad=True)
    mu = A.solute_mu(p, names, W, s, chi_w)
    mu.backward()
    g = p.grad.numpy()
    h = 1e-4
    for a, c in ((0, 1), (2, 3), (3, 4), (1, 5)):
        d = np.zeros_like(p0)
        d[a, c] = h
        with torch.no_grad():
            fp = float(A.solute_mu(torch.tensor(p0 + d), names, W, s, chi_w))
            fm = float(A.solute_mu(torch.tensor(p0 - d), names, W, s, chi_w))
        fd = (fp - fm) / (2 * h)
        print(f"  site {a} {A.COLUMNS[c]}: autograd {g[a, c]:.8f} fd {fd:.8f}", flush=True)
        assert abs(g[a, c] - fd) <= 1e-4 * max(abs(fd), 1e-2)
