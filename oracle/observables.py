"""Post-solve observables of the solute path (oracle): `info`.

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates info / print_info (bgy3d-potential.c:340-491) with cage = false
(hnc3d.c:1774): the solvent charge density rho_v = rho sum_i q_i g_i (g as
hnc3d.c:2319-2325 passes it), its potential from the FFT Poisson solver
(bgy3d_poisson, bgy3d-poisson.c:103-190, k^2 summed z, y, x), the Gaussian
solute cores rho_u (bgy3d-solutes.c:555-577), the moments of both
(bgy3d_moments), the solvent potential at the solute sites by
trigonometric interpolation (bgy3d_fft_interp, bgy3d-fftw3.c:504-612) and
the three interaction integrals.  Pinned by the lines test/summary.sh
extracts from test/out/*.hnc3d.g1.
"""
import math

import numpy as np

from . import rism3d as O
from .solvent3d import moments

EPS0INV = O.EPSILON0INV
G = O.G


def solvent_charge_density(pd, solvent, g):
    return pd.rho * sum(s[4] * g[i] for i, s in enumerate(solvent))


def poisson(pd, rho):
    """Potential of rho in kcal (q = -4 pi / eps0), k = 0 term zero."""
    kx, ky, kz = O.k_axes(pd)
    k2 = (kz[:, None, None] ** 2 + ky[None, :, None] ** 2) + kx[None, None, :] ** 2
    scale = 4 * math.pi * EPS0INV * pd.h3
    with np.errstate(divide="ignore"):
        fac = np.where(k2 != 0.0, scale / k2, 0.0)
    return O.irfft(pd, O.rfft(pd, rho) * fac) / pd.L3


def cores(pd, solute):
    x, y, z = O.r_axes(pd)
    acc = np.zeros(pd.N[::-1])
    for s in solute:
        r2 = ((x[None, None, :] - s[1][0]) ** 2 + (y[None, :, None] - s[1][1]) ** 2) \
            + (z[:, None, None] - s[1][2]) ** 2
        acc += s[4] * np.exp(-G * G * r2)
    return (G / math.sqrt(math.pi)) ** 3 * acc


def fft_interp(pd, v, xs):
    """Trigonometric interpolation of the real field v at points xs [A]."""
    N0, N1, N2 = pd.N
    Y = O.rfft(pd, v)
    out = []
    for p in xs:
        y = [(p[d] + pd.L[d] / 2) / pd.h[d] for d in range(3)]

        def wave(n, ks, coord):
            K = np.array([k if k <= n // 2 else k - n for k in ks])
            arg = (2 * math.pi / n) * K * coord
            real_only = ((n - 2 * np.asarray(ks)) % n) == 0
            return np.where(real_only, np.cos(arg) + 0j, np.exp(1j * arg))

        wx = wave(N0, range(N0 // 2 + 1), y[0])
        wy = wave(N1, range(N1), y[1])
        wz = wave(N2, range(N2), y[2])
        w = wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        half = np.where(((N0 - 2 * np.arange(N0 // 2 + 1)) % N0) == 0, 0.5, 1.0)
        out.append(2 * float(np.sum(np.real(Y * w * half[None, None, :]))) / (N0 * N1 * N2))
    return np.array(out)


def info(pd, solvent, solute, g):
    """The numbers print_info prints, as a dict."""
    rho_v = solvent_charge_density(pd, solvent, g)
    ve = poisson(pd, rho_v)
    rho_u = cores(pd, solute)
    uc = O.coulomb_long(pd, solute)
    d_n = np.sum([np.multiply(s[4], s[1]) for s in solute], axis=0)
    q_u, d_u, _ = moments(pd, rho_u)
    q_v, d_v, _ = moments(pd, rho_v)
    vs = fft_interp(pd, ve, [s[1] for s in solute])
    return {
        "q_v": q_v, "q_u": q_u,
        "d_n": list(map(float, d_n)), "d_u": d_u, "d_v": d_v,
        "U_v.rho_N": float(sum(s[4] * v for s, v in zip(solute, vs))),
        "U_v.rho_u": pd.h3 * float(np.sum(ve * rho_u)),
        "rho_v.U_u": pd.h3 * float(np.sum(uc * rho_v)),
        "v_at_sites": vs,
    }
