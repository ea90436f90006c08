"""The oracle's closure continuation (KH, PSE-2, ..., then HNC) for a
solute whose HNC iteration diverges from t = 0 with the default MDIIS step:
it must land on the same root as a direct MDIIS run with a smaller step
(which converges slowly but without the ladder)."""
import numpy as np
import pytest

from conftest import CHI_W, load_chi
from oracle import closures as C
from oracle import rism3d as O
from oracle.nksolve import NotConverged, mdiis
from bgy3d_amd import synthetic as S
from bgy3d_amd.solutes import find_molecule

W = find_molecule("water, PR-SPC/E")


def test_ladder_reaches_the_direct_root():
    chi = load_chi(CHI_W)
    sol = list(S.solute(4, n=6))
    pd = O.Problem.from_cli(8.0, 32, 1.6889, 0.0333295, closure=C.HNC, rmax=40.0, nrad=1536)
    s = O.Solute3D(pd, W, sol, chi)
    with pytest.raises(NotConverged, match="not finite"):
        mdiis(s.iterate, np.zeros_like(s.v), atol=1e-10, lam=0.5, depth=10)
    assert s.ladder() == [C.KH, C.PSE0 + 2, C.PSE0 + 3, C.PSE0 + 4]
    t, _ = s.solve(atol=1e-10)
    assert pd.closure == C.HNC
    assert np.max(np.abs(s.iterate(t))) <= 1e-10
    t2, _ = mdiis(s.iterate, np.zeros_like(s.v), atol=1e-10, lam=0.2, depth=8, max_iter=3000)
    assert np.max(np.abs(t - t2)) < 1e-7
    assert abs(s.observables(t)["mu"] - s.observables(t2)["mu"]) < 1e-8
