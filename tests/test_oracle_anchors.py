"""Pin the oracle end to end against the reference's published numbers.

README:9-16   1D DRISM + 1D solute RISM, PR-SPC/E water in itself,
              rmax 160, nrad 4096, eps 78.4: mu = -7.869294
README:20-31  3D-RISM, same molecules, --L 10 --N 96: mu = -6.349052
              (renorm), -6.297170 (--no-renorm).  chi from 1D DRISM on the
              grid hnc3d_solute_solve builds itself (rmax 40, nrad 1536).
(The net-charged rows of test/out/run-ions-out are pinned in
test_oracle_run_ions.py.)
Also: a cyclic permutation of the axes (an exact symmetry of a cubic grid
with the same N and L on every axis) leaves mu unchanged for an uncharged
solute; a transposed index anywhere in the grid code would break it.  (With
charged sites tau F(k) is not Hermitian on the Nyquist planes and FFTW's c2r
drops Im only on the half-complex x axis, so the axes are not equivalent
there -- reference behaviour, reproduced by the GPU path.)
"""
import numpy as np

from conftest import CHI_W, load_chi
from oracle import closures as C
from oracle import rism1d, rism3d
from bgy3d_amd.solutes import find_molecule

W = find_molecule("water, PR-SPC/E")


def test_readme_1d():
    vv = rism1d.rism_vv(C.HNC, 4096, 160.0, 1.6889, 0.0333295, W, eps=78.4)
    uv = rism1d.rism_uv(C.HNC, 4096, 160.0, 1.6889, 0.0333295, W, vv["chi"], W)
    assert abs(uv["mu"] - (-7.869294)) < 1e-6


def test_fixture_chi_is_oracle_chi():
    vv = rism1d.rism_vv(C.HNC, 1536, 40.0, 1.6889, 0.0333295, W, eps=78.4)
    np.testing.assert_allclose(vv["chi"], load_chi(CHI_W), atol=1e-10)


def test_readme_3d_renorm():
    chi = load_chi(CHI_W)
    pd = rism3d.Problem.from_cli(10.0, 96, 1.6889, 0.0333295, rmax=40.0, nrad=1536)
    out = rism3d.solute_solve(pd, W, W, chi, atol=1e-10)
    assert abs(out["mu"] - (-6.349052)) < 1e-6


def test_readme_3d_norenorm():
    chi = load_chi(CHI_W)
    pd = rism3d.Problem.from_cli(10.0, 96, 1.6889, 0.0333295, rmax=40.0, nrad=1536)
    out = rism3d.solute_solve(pd, W, W, chi, renorm=False, atol=1e-10)
    assert abs(out["mu"] - (-6.297170)) < 1e-6


def test_axis_permutation_symmetry():
    # rmax 39.7: the table spacing pi/rmax is incommensurate with 2 pi/L, so
    # no grid |k| sits on a midpoint of interp's table (see oracle.kmag);
    # any smooth chi serves for a symmetry check.
    chi = load_chi(CHI_W)
    pd = rism3d.Problem.from_cli(10.0, 32, 1.6889, 0.0333295, rmax=39.7, nrad=1536)
    ow2 = [(n, (0.3, -0.2, x[2]), s, e, q) for n, x, s, e, q in find_molecule("OW2")]
    rot = [(n, (x[2], x[0], x[1]), s, e, q) for n, x, s, e, q in ow2]
    a = rism3d.solute_solve(pd, W, ow2, chi, atol=1e-11)
    b = rism3d.solute_solve(pd, W, rot, chi, atol=1e-11)
    assert abs(a["mu"] - b["mu"]) < 1e-9
    # g permutes with the coordinates: g'(z, y, x) = g(y, x, z)
    np.testing.assert_allclose(b["g"], np.transpose(a["g"], (0, 2, 3, 1)), atol=1e-8)


def test_forces_scm_anchors():
    """test/forces.scm:28-30: HNC, --L 10 --N 96, OW solvent, chi from the
    1D run hnc3d_solute_solve makes itself (upscaled rmax 40, nrad 1536,
    plain RISM): mu(OW2, d/2 = 2) = 14.0074968008389, mu(OW) =
    6.86940765594207, mu(OW2-UA) = 5.82278007673305."""
    ow = find_molecule("OW")
    vv = rism1d.rism_vv(C.HNC, 1536, 40.0, 1.6889, 0.0333295, ow)
    pd = rism3d.Problem.from_cli(10.0, 96, 1.6889, 0.0333295, rmax=40.0, nrad=1536)
    for name, ref in (("OW", 6.86940765594207), ("OW2-UA", 5.82278007673305),
                      ("OW2", 14.0074968008389)):
        out = rism3d.solute_solve(pd, ow, find_molecule(name), vv["chi"], atol=1e-12)
        assert abs(out["mu"] - ref) < 1e-8, (name, out["mu"])
