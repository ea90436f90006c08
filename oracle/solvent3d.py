"""3D solvent-solvent HNC/KH (oracle), the `--solvent-3d` susceptibility.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates hnc3d_solvent_solve (hnc3d.c:804-1073): site-site pair functions
t_ij(r) of a molecular solvent on the 3D grid, centred at the box centre,
solved from the OZ equation in k-space with the intramolecular omega(k):

* pair potential (bgy3d_pair_potential, bgy3d-force.c:123-143): short range
  LJ + erfc Coulomb with the reference's caps (bgy3d-force.h:23-109),
  long range 4 pi/eps0 q_i q_j exp(-k^2/4G^2)/k^2 tabulated around the
  corner and translated to the centre (bgy3d-force.c:19-36);
* omega_ij(k) = sinc(k r_ij) (bgy3d-pure.c:236-313), corner;
* residual (iterate_t2, hnc3d.c:376-470): C = h^3 FFT(c) - beta V_L,
  translated to the corner; per k, H = (1 - rho W C)^-1 W C W and
  T = H - C (compute_t2_m, hnc3d.c:231-324, solved here by numpy instead of
  LAPACK); T translated back, minus beta V_L, inverse FFT / L^3, minus t;
* post: h = c + t, mu by the m x m chemical potential functional
  (print_chempot, hnc3d.c:669-687, cl = -beta IFFT(V_L)/L^3), and
  chi - 1 = FFT(rho h) h^3 at the corner + (omega - 1), the array the
  reference saves as x%d%d-fft.bin (hnc3d.c:1006-1052).

Pinned by test/out/water.hnc3d.g2 (tests/test_oracle_solvent3d.py): the
three mu functionals and the moments of 1 - g_ij.
"""
import math

import numpy as np

from . import closures as C
from . import rism3d as O
from .nksolve import mdiis

EPS0INV = O.EPSILON0INV
G = O.G
CUTOFF = 1.0e8
CUTOFF2 = CUTOFF * 1.0e-5


def lennard_jones(r, eps, sig):
    """bgy3d-force.h:23-37 (uncapped form with the CUTOFF ceiling)."""
    if sig == 0.0 or eps == 0.0:
        return np.zeros_like(r)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        sr = sig / r
        sr6 = (sr * sr) * (sr * sr) * (sr * sr)
        re = 4 * eps * sr6 * (sr6 - 1)
    return np.where(np.abs(re) > eps * CUTOFF, eps * CUTOFF, re)


def coulomb_short(r, q2, g=G):
    """bgy3d-force.h:76-88."""
    from scipy.special import erfc
    with np.errstate(divide="ignore", invalid="ignore"):
        f = erfc(g * r) / r
    return np.where(f > CUTOFF2, EPS0INV * q2 * CUTOFF2, EPS0INV * q2 * f)


def r_center(pd):
    x, y, z = O.r_axes(pd)
    return np.sqrt((x[None, None, :] ** 2 + y[None, :, None] ** 2) + z[:, None, None] ** 2)


def coulomb_long_fft_unit(pd, g=G):
    """bgy3d_coulomb_long_fft (bgy3d-force.c:19-36): vec_kmap takes
    k = sqrt(kx^2 + ky^2 + kz^2) and coulomb_long_fourier squares it back."""
    k = O.kmag(pd)
    k2 = k * k
    with np.errstate(divide="ignore", invalid="ignore"):
        f = (4 * math.pi * EPS0INV) * np.exp(-k2 / (4 * g * g)) / k2
    f = np.where(k2 == 0.0, 0.0, f)
    return f * O.phase_sign(pd)


def pairs(m):
    return [(i, j) for i in range(m) for j in range(i + 1)]


def omega_fft(pd, solvent):
    """omega_ij(k) = sinc(2 pi r_ij / L sqrt(k2)), k2 the integer sum
    kz^2 + ky^2 + kx^2 (bgy3d-pure.c:236-267)."""
    N0, N1, N2 = pd.N
    ix = np.arange(N0 // 2 + 1)
    iy = O.kfreq(N1)
    iz = O.kfreq(N2)
    k2 = (iz[:, None, None] ** 2 + iy[None, :, None] ** 2) + ix[None, None, :] ** 2
    m = len(solvent)
    w = {}
    for i, j in pairs(m):
        if i == j:
            continue
        r = float(np.linalg.norm(np.subtract(solvent[i][1], solvent[j][1])))
        kr = (2.0 * math.pi * r / pd.L[0]) * np.sqrt(k2.astype(np.float64))
        with np.errstate(divide="ignore", invalid="ignore"):
            w[(i, j)] = np.where(kr == 0.0, 1.0, np.sin(kr) / np.where(kr == 0.0, 1.0, kr))
    return w


class Solvent3D:
    """Everything hnc3d_solvent_solve sets up before SNESSolve."""

    def __init__(self, pd, solvent):
        self.pd = pd
        self.solvent = solvent
        self.m = len(solvent)
        self.pairs = pairs(self.m)
        r = r_center(pd)
        self.v = np.stack([
            lennard_jones(r, math.sqrt(solvent[i][3] * solvent[j][3]),
                          0.5 * (solvent[i][2] + solvent[j][2]))
            + coulomb_short(r, solvent[i][4] * solvent[j][4])
            for i, j in self.pairs])
        ul = coulomb_long_fft_unit(pd)
        self.vl = np.stack([solvent[i][4] * solvent[j][4] * ul for i, j in self.pairs])
        self.w = omega_fft(pd, solvent)
        self.s = O.phase_sign(pd)

    def _matrix(self, vals):
        """[np][...] pair values -> [..., m, m] symmetric."""
        m = self.m
        shape = vals.shape[1:]
        M = np.zeros(shape + (m, m), dtype=vals.dtype)
        for p, (i, j) in enumerate(self.pairs):
            M[..., i, j] = vals[p]
            M[..., j, i] = vals[p]
        return M

    def iterate(self, t):
        pd = self.pd
        beta, rho = pd.beta, pd.rho
        c = C.closure(pd.closure, beta, self.v, t)
        Cf = np.stack([(O.rfft(pd, c[p]) * pd.h3 - beta * self.vl[p]) * self.s
                       for p in range(len(self.pairs))])
        Cm = self._matrix(Cf)
        Wm = np.zeros(Cm.shape)
        for i in range(self.m):
            Wm[..., i, i] = 1.0
        for (i, j), w in self.w.items():
            Wm[..., i, j] = w
            Wm[..., j, i] = w
        WC = Wm @ Cm
        WCW = WC @ Wm
        A = np.eye(self.m) - rho * WC
        H = np.linalg.solve(A, WCW)
        out = []
        for p, (i, j) in enumerate(self.pairs):
            Tp = (H[..., i, j] - Cm[..., i, j]) * self.s - beta * self.vl[p]
            out.append(O.irfft(pd, Tp) / pd.L3)
        return np.stack(out) - t

    def initial_guess(self, t_rad, rmax):
        """1D t_ij(r) tabulated around the centre (vec_rtab,
        bgy3d-vec.c:49-57); t_rad[m][m][nrad]."""
        nrad = t_rad.shape[-1]
        dr = rmax / nrad
        r = r_center(self.pd)
        return np.stack([O.interp(r, t_rad[i][j], dr) for i, j in self.pairs])

    def solve(self, t0=None, atol=1e-9, **kw):
        t0 = np.zeros_like(self.v) if t0 is None else t0
        t, it = mdiis(self.iterate, t0, atol=atol, **kw)
        return t, it

    def observables(self, t):
        pd = self.pd
        beta = pd.beta
        c = C.closure(pd.closure, beta, self.v, t)
        h = c + t
        x = -beta * self.v + t
        cl = np.stack([-beta * O.irfft(pd, self.vl[p]) / pd.L3 for p in range(len(self.pairs))])
        mult = np.array([1.0 if i == j else 2.0 for i, j in self.pairs])

        def mu_of(method):
            dens = C.form(method, x, h, c, cl)
            tot = sum(mult[p] * math.fsum(dens[p].ravel()) for p in range(len(self.pairs)))
            return pd.rho * tot * pd.h3 / beta

        chi = {}
        for p, (i, j) in enumerate(self.pairs):
            f = O.rfft(pd, pd.rho * h[p]) * pd.h3 * self.s
            if i != j:
                f = f + self.w[(i, j)]
            chi[(i, j)] = f
        return {"mu_hnc": mu_of(C.HNC), "mu_kh": mu_of(C.KH), "mu_gf": mu_of(C.PY),
                "g": h + 1.0, "chi": chi}


def moments(pd, v):
    """bgy3d_moments (bgy3d-vec.c:573-593): charge, dipole and second
    moments about the grid centre, index offsets i - N/2."""
    N0, N1, N2 = pd.N
    h0, h1, h2 = pd.h
    dV = h0 * h1 * h2
    ix = np.arange(N0) - 0.5 * N0
    iy = np.arange(N1) - 0.5 * N1
    iz = np.arange(N2) - 0.5 * N2
    X = [ix[None, None, :], iy[None, :, None], iz[:, None, None]]
    hh = [h0, h1, h2]
    q = dV * float(np.sum(v))
    d = [dV * hh[a] * float(np.sum(v * X[a])) for a in range(3)]
    Q = [[dV * hh[a] * hh[b] * float(np.sum(v * X[a] * X[b])) for b in range(3)]
         for a in range(3)]
    return q, d, Q


def chi_table(obs, m):
    """chi - 1 of the 3D solvent as the real [m(m+1)/2][N2][N1][N0/2+1]
    table the solute kernel takes (pair p = i(i+1)/2 + j, j <= i); the
    imaginary part vanishes for the inversion-symmetric solution."""
    return np.stack([obs["chi"][(i, j)].real for i, j in pairs(m)])
