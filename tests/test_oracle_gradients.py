"""Pin oracle/gradients.py: the TPT field derivative (field1's ljcap1 and
ccap1, bgy3d-solutes.c:381-468) against central differences of the full
capped site potential, the adjoint d mu / d theta against central
differences of converged mu, all six parameter kinds, and the reference's
linear-response gradient (hnc3d.c:1366-1421, 1969-2134) against both: equal
to the exact gradient where the solute is uncharged, equal to the one-solve
adjoint form for a charged one."""
import numpy as np
import pytest

from conftest import CHI_W, load_chi
from oracle import gradients as Gr
from oracle import rism3d as O
from bgy3d_amd.solutes import find_molecule
from bgy3d_amd import synthetic as S

W = find_molecule("water, PR-SPC/E")


def test_capped_derivatives_fd():
    r = np.linspace(0.01, 3.0, 400)
    r = r[np.abs(r - 0.2) > 1e-3]
    h = 1e-6
    for f0, f1 in ((O.ljcap0, Gr.ljcap1), (O.cscap0, Gr.cscap1), (O.cl0, Gr.cl1), (Gr.ccap0, Gr.ccap1)):
        fd = (f0(r + h) - f0(r - h)) / (2 * h)
        np.testing.assert_allclose(f1(r), fd, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize("renorm", [True, False])
def test_adjoint_matches_fd(renorm):
    """Both paths: tau renormalisation and the Ng form (--no-renorm), whose
    long-range weight the adjoint builds from chi and u_c(k)."""
    chi = load_chi(CHI_W)
    sol = list(S.solute(5, n=4))
    pd = O.Problem.from_cli(8.0, 32, 1.6889, 0.0333295, rmax=40.0, nrad=1536)

    def mu(sites):
        s = O.Solute3D(pd, W, sites, chi, renorm=renorm)
        t, _ = s.solve(atol=1e-12)
        return s.observables(t)["mu"], s, t

    _, s0, t0 = mu(sol)
    grad, _ = Gr.adjoint_gradient(s0, t0)
    h = 1e-4
    for a, p in ((0, 1), (1, 3), (2, 4), (3, 5)):
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
        fd = (mu(moved(+1))[0] - mu(moved(-1))[0]) / (2 * h)
        assert abs(grad[a, p] - fd) <= 1e-5 * max(abs(fd), 1e-2), (a, p, grad[a, p], fd)


def _solved(pd, sites, chi, solvent=W):
    s = O.Solute3D(pd, solvent, sites, chi)
    t, _ = s.solve(atol=1e-12)
    return s, t


def test_response_uncharged_equals_exact_gradient():
    """For an uncharged solute the reference's linear-response gradient
    (hnc3d.c:1969-2134) leaves nothing out: it equals the adjoint
    gradient's x, y, z columns, central differences of mu, and the TPT
    gradient (hnc3d.c:1666-1742), exact here as the HNC mu is stationary."""
    chi = load_chi(CHI_W)
    pd = O.Problem.from_cli(8.0, 32, 1.6889, 0.0333295, rmax=40.0, nrad=1536)
    sites = [("A", (0.0, 0.0, 0.9), 3.16, 0.155, 0.0), ("B", (0.3, 0.1, -0.9), 3.0, 0.1, 0.0)]
    s, t = _solved(pd, sites, chi)
    resp = Gr.response_gradient(s, t)
    grad, _ = Gr.adjoint_gradient(s, t)
    np.testing.assert_allclose(resp, grad[:, :3], rtol=0, atol=1e-6 * np.max(np.abs(resp)))
    # and so is the TPT form: uncharged, the mu functional is stationary
    tpt = Gr.tpt_gradient(pd, W, sites, s.observables(t)["g"])
    np.testing.assert_allclose(tpt, grad[:, :3], rtol=0, atol=1e-6 * np.max(np.abs(resp)))
    h = 1e-4
    for a, d in ((0, 2), (1, 0)):
        mus = []
        for sign in (+1, -1):
            moved = [list(x) for x in sites]
            xyz = list(moved[a][1])
            xyz[d] += sign * h
            moved[a][1] = tuple(xyz)
            ss, tt = _solved(pd, [tuple(x) for x in moved], chi)
            mus.append(ss.observables(tt)["mu"])
        fd = (mus[0] - mus[1]) / (2 * h)
        assert abs(resp[a, d] - fd) <= 1e-5 * max(abs(fd), 1e-2), (a, d, resp[a, d], fd)


def test_response_literal_equals_adjoint_form():
    """Charged solute, KH: the 3n linear solves of the reference's response
    path and the one-solve adjoint form give the same gradient."""
    chi = load_chi(CHI_W)
    pd = O.Problem.from_cli(8.0, 32, 1.6889, 0.0333295, closure=1, rmax=40.0, nrad=1536)
    sites = find_molecule("water, cSPC/E")
    s, t = _solved(pd, sites, chi)
    lit = Gr.response_gradient(s, t)
    lam = Gr.adjoint_lambda(s, t)
    adj = Gr.response_from_adjoint(s, t, lam)
    np.testing.assert_allclose(lit, adj, rtol=0, atol=1e-7 * np.max(np.abs(lit)))
    # the long-range terms the response path leaves out are not small here
    grad, _ = Gr.adjoint_gradient(s, t)
    assert np.max(np.abs(grad[:, :3] - lit)) > 1e-3 * np.max(np.abs(lit))
