"""1D RISM / DRISM solvent susceptibility and 1D solute RISM (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates the reference's chi producer and its 1D solute counterpart:

* mkgrid            fft.f90:101-118   r_i=(i-1/2)dr, k_i=(i-1/2)dk, dk=pi/rmax
* fourier_rows      fft.f90:121-170   radial FT via DST-IV (FFTW RODFT11 ==
                                      scipy.fft.dst(type=4), unnormalised)
* integrate         fft.f90:242-267   4 pi sum f_i (2i-1)^2/4 (dr = 1)
* force_field       rism.f90:2082-2162 LJ + erfc short range on r, Coulomb
                                      long range 4 pi e^{-k^2/4a^2}/k^2 on k
* force_field_long  rism.f90:2165-2198
* omega_fourier     rism.f90:1994-2032 j0(k r_ab)
* j0, j1            bessel.f90:36-91
* dipole_correction drism.f90:380-474 (+ dipole_axes 223-354, center 148)
* oz_vv_equation_c_t_MxM rism.f90:2476-2560, oz_uv_equation_c_t 2583-2607
* rism_vv           rism.f90:310-583  (solve for r*t, chi = w + rho h)
* rism_uv           rism.f90:586-859  (renormalised Ng form, s_uvk)
* rism_solvent_renorm rism.f90:141-208
* chempot           closures.f90:790-829

Sites are sequences ``(name, (x, y, z), sigma, epsilon, charge)``.

Pinning: README:9-16 (1D mu = -7.869294 for PR-SPC/E water in itself,
rmax 160, nrad 4096, eps 78.4), and end-to-end through the 3D README anchor
(-6.349052), in tests/test_oracle_anchors.py; DST-IV against an explicit
O(n^2) sine sum and the Gaussian FT closed form (fft.f90:368-432) in
tests/test_oracle_rism1d.py.
"""
import math

import numpy as np
import scipy.fft

from . import closures as C
from .nksolve import newton_krylov

EPSILON0INV = 331.84164   # bgy3d.h:145, units.f90:50
ALPHA = 1.2               # units.f90:61, bgy3d-solvents.h:9
PI = math.pi


def mkgrid(rmax, nrad):
    dr = rmax / nrad
    dk = PI / rmax
    i = np.arange(1, nrad + 1, dtype=np.float64)
    return (2 * i - 1) * dr / 2, dr, (2 * i - 1) * dk / 2, dk


def dst4(a, axis=-1):
    """FFTW RODFT11: y_k = 2 sum_j x_j sin(pi (2j+1)(2k+1) / 4n)."""
    return scipy.fft.dst(a, type=4, axis=axis)


def fourier_rows(f):
    """fft.f90:121-170; f has the radial index last."""
    n = f.shape[-1]
    p = np.arange(1, n + 1, dtype=np.float64)
    g = dst4(f * (2 * n * (2 * p - 1)))
    return g / (2 * p - 1)


def integrate(f):
    """fft.f90:242-267: 4 pi sum_i f_i (2i-1)^2/4, summed backwards."""
    n = f.shape[-1]
    i = np.arange(1, n + 1, dtype=np.float64)
    w = (2 * i - 1) ** 2 / 4
    return 4 * PI * math.fsum((f * w)[::-1])


def j0(x):
    x = np.asarray(x, dtype=np.float64)
    y = x * x
    small = 1 + y * (-1 / 6.0 + y * (1 / 120.0 + y * (-1 / 5040.0 + y * (
        1 / 362880.0 + y * (-1 / 39916800.0 + y * (1 / 6227020800.0))))))
    with np.errstate(invalid="ignore", divide="ignore"):
        big = np.sin(x) / x
    return np.where(np.abs(x) < 0.5, small, big)


def j1(x):
    x = np.asarray(x, dtype=np.float64)
    y = x * x
    s = 1 + y * (-1 / 10.0 + y * (1 / 280.0 + y * (-1 / 15120.0 + y * (
        1 / 1330560.0 + y * (-1 / 172972800.0)))))
    with np.errstate(invalid="ignore", divide="ignore"):
        big = (np.sin(x) / x - np.cos(x)) / x
    return np.where(np.abs(x) < 0.25, x / 3 * s, big)


def lj(r):
    sr6 = 1.0 / r ** 6
    return 4 * sr6 * (sr6 - 1)


def coulomb_long_fourier(k, alpha=ALPHA):
    k = np.asarray(k, dtype=np.float64)
    with np.errstate(divide="ignore"):
        f = 4 * PI * np.exp(-k ** 2 / (4 * alpha ** 2)) / k ** 2
    return np.where(k == 0.0, 0.0, f)


def coulomb_long(r, alpha=ALPHA):
    from scipy.special import erf
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = erf(alpha * r) / r
    return np.where(r == 0.0, alpha * 2 / math.sqrt(PI), f)


def coulomb_short(r, alpha=ALPHA):
    from scipy.special import erfc
    return erfc(alpha * r) / r


def _xyz(s):
    return np.asarray(s[1], dtype=np.float64)


def force_field(asites, bsites, r, k):
    """rism.f90:2082-2162 with the Lorentz-Berthelot rule (rism.f90:1323)."""
    n, m = len(asites), len(bsites)
    vr = np.zeros((n, m, r.size))
    vk = np.zeros((n, m, k.size))
    for j, b in enumerate(bsites):
        for i, a in enumerate(asites):
            qab = a[4] * b[4]
            sigma = (a[2] + b[2]) / 2
            eps = math.sqrt(a[3] * b[3])
            if sigma != 0.0:
                vr[i, j] = eps * lj(r / sigma) + EPSILON0INV * qab * coulomb_short(r)
            else:
                vr[i, j] = EPSILON0INV * qab * coulomb_short(r)
            vk[i, j] = EPSILON0INV * qab * coulomb_long_fourier(k)
    return vr, vk


def force_field_long(solute, solvent, r):
    v = np.zeros((len(solute), len(solvent), r.size))
    cl = coulomb_long(r)
    for i, a in enumerate(solute):
        for j, b in enumerate(solvent):
            v[i, j] = a[4] * b[4] * EPSILON0INV * cl
    return v


def omega_fourier(sites, k, molecule=None):
    """w_ij(k) = j0(k r_ij); with `molecule` (a molecule index per site)
    zero between sites of different molecules (the intramolecular matrix of
    a mixture)."""
    m = len(sites)
    w = np.zeros((m, m, k.size))
    for j in range(m):
        for i in range(m):
            if molecule is not None and molecule[i] != molecule[j]:
                continue
            rab = np.linalg.norm(_xyz(sites[i]) - _xyz(sites[j]))
            w[i, j] = j0(k * rab)
    return w


# ---- DRISM (drism.f90) ------------------------------------------------

def dipole(sites):
    return sum(s[4] * _xyz(s) for s in sites)


def dipole_density(beta, rho, sites):
    return 4 * PI * beta * EPSILON0INV * rho * np.linalg.norm(dipole(sites)) ** 2 / 9


def epsilon_rism(beta, rho, sites):
    return 1 + 3 * dipole_density(beta, rho, sites)


def _weights(sites):
    w = np.array([s[2] ** 3 for s in sites], dtype=np.float64)
    return w / w.sum()


def center(sites):
    w = _weights(sites)
    return sum(wi * _xyz(s) for wi, s in zip(w, sites))


def _axes(sites):
    c = center(sites)
    w = _weights(sites)
    t = np.zeros((3, 3))
    for wi, s in zip(w, sites):
        d = _xyz(s) - c
        t += wi * np.outer(d, d)
    _e, u = np.linalg.eigh(t)
    return u


def dipole_axes(sites):
    """drism.f90:223-354.  x and y are the eigenvectors of the 2x2 shape
    tensor in the plane orthogonal to the dipole; only |x_a|, |y_a| and z_a
    enter dipole_correction, so sign/order of the eigenvectors is
    immaterial."""
    mu = dipole(sites)
    mu = mu / np.linalg.norm(mu)
    v = _axes(sites)
    e = np.zeros(3)
    for j in range(3):
        v[:, j] = v[:, j] - mu * np.dot(mu, v[:, j])
        e[j] = np.dot(v[:, j], v[:, j])
    i = int(np.argmin(e))
    u = np.zeros((3, 3))
    kk = 0
    for j in range(3):
        if j == i:
            continue
        u[:, kk] = v[:, j]
        kk += 1
    u[:, 2] = mu
    u[:, 1] = u[:, 1] - u[:, 0] * np.dot(u[:, 0], u[:, 1]) / np.dot(u[:, 0], u[:, 0])
    for j in range(3):
        u[:, j] = u[:, j] / np.linalg.norm(u[:, j])
    c = center(sites)
    w = _weights(sites)
    t = np.zeros((2, 2))
    for wi, s in zip(w, sites):
        d = _xyz(s) - c
        dxy = np.array([np.dot(d, u[:, 0]), np.dot(d, u[:, 1])])
        t += wi * np.outer(dxy, dxy)
    _e2, v2 = np.linalg.eigh(t)
    u[:, 0:2] = u[:, 0:2] @ v2
    return u


def local_coords(sites, u):
    c = center(sites)
    return np.array([[np.dot(u[:, i], _xyz(s) - c) for i in range(3)]
                     for s in sites]).T          # (3, m)


def dipole_correction(beta, rho, eps, sites, k):
    """drism.f90:380-474 (s = 0.5 A, Ref. KH00a)."""
    s = 0.5
    x = local_coords(sites, dipole_axes(sites))
    m = len(sites)
    f = np.zeros((k.size, m))
    for i in range(m):
        d = x[:, i]
        f[:, i] = j0(k * d[0]) * j0(k * d[1]) * j1(k * d[2])
    y = dipole_density(beta, rho, sites)
    h = (eps - (1 + 3 * y)) / (rho * y) * np.exp(-(k * s) ** 2 / 4)
    xk = np.zeros((m, m, k.size))
    for j in range(m):
        for i in range(m):
            xk[i, j] = f[:, i] * h * f[:, j]
    return xk


# ---- OZ equations -----------------------------------------------------

def oz_vv_equation_c_t(rho, c, w):
    """rism.f90:2432-2560; arrays (m, m, nrad)."""
    m = c.shape[0]
    if m == 1:
        cc = c[0, 0]
        return (rho * (cc * cc) / (1 - rho * cc))[None, None, :]
    ck = np.moveaxis(c, -1, 0)
    wk = np.moveaxis(w, -1, 0)
    H = wk @ ck
    T = np.eye(m)[None] - H * rho  # rho: scalar, or per site (scales column j)
    H = H @ wk
    H = np.linalg.solve(T, H)
    return np.moveaxis(H - ck, 0, -1)


def oz_vv_equation_c_t1(rho, c, w, dc):
    """Differential of oz_vv_equation_c_t with respect to c.  From
    A H = W C W with A = 1 - W C rho:
    dH = A^-1 W dC (W + rho H), dT = dH - dC.  (The reference leaves this
    Jacobian to PETSc finite differences; only the root matters.)"""
    m = c.shape[0]
    if m == 1:
        cc = c[0, 0]
        return (rho * cc * (2 - rho * cc) / (1 - rho * cc) ** 2 * dc[0, 0])[None, None, :]
    ck = np.moveaxis(c, -1, 0)
    wk = np.moveaxis(w, -1, 0)
    dck = np.moveaxis(dc, -1, 0)
    A = np.eye(m)[None] - (wk @ ck) * rho
    H = np.linalg.solve(A, wk @ ck @ wk)
    rho_rows = np.asarray(rho)[..., None] if np.ndim(rho) else rho  # (rho H)_qj = rho_q H_qj
    dH = np.linalg.solve(A, wk @ dck @ (wk + rho_rows * H))
    return np.moveaxis(dH - dck, 0, -1)


def oz_uv_equation_c_t(c_uv, w_uu, x_vv):
    """rism.f90:2583-2607: t = w * c * x - c per k."""
    c = np.moveaxis(c_uv, -1, 0)
    w = np.moveaxis(w_uu, -1, 0)
    x = np.moveaxis(x_vv, -1, 0)
    return np.moveaxis(w @ (c @ x) - c, 0, -1)


def chempot(method, rmax, beta, rho, v, vl, t):
    """closures.f90:790-829."""
    nrad = t.shape[-1]
    _r, dr, _k, _dk = mkgrid(rmax, nrad)
    c = C.closure(method, beta, v, t)
    cl = -beta * vl
    h = c + t
    x = -beta * v + t
    dens = C.form(method, x, h, c, cl).sum(axis=(0, 1))
    return integrate(dens) * (rho * dr ** 3 / beta)


def rism_vv(method, nrad, rmax, beta, rho, sites, eps=0.0, atol=1e-12,
            verbose=False):
    """rism.f90:310-583.  Returns dict(t, chi, c, mu, stats).  chi is
    (m, m, nrad) in Fortran (m, m, nrad) order; the C buffer layout
    double[m][m][nrad] of rism_solvent (rism.f90:89-96) is the same array
    read in C order because chi is symmetric."""
    m = len(sites)
    r, dr, k, dk = mkgrid(rmax, nrad)
    vr, vk = force_field(sites, sites, r, k)
    wk = omega_fourier(sites, k)
    xk = dipole_correction(beta, rho, eps, sites, k) if eps != 0.0 else np.zeros_like(wk)
    A = 1.0
    FT_BW = (2 * PI) ** 3

    def iterate_t(t):
        c = C.closure(method, beta, vr, t)
        c = fourier_rows(c) * dr ** 3
        c = c - (beta * A) * vk
        if eps != 0.0:
            dt = oz_vv_equation_c_t(rho, c, wk + rho * xk) + xk
        else:
            dt = oz_vv_equation_c_t(rho, c, wk)
        dt = dt - (beta * A) * vk
        dt = fourier_rows(dt) * (dk ** 3 / FT_BW)
        return dt - t

    def iterate_rt(rt):
        return r * iterate_t(rt / r)

    W = wk + rho * xk

    def jacobian_rt(rt, drt):
        t = rt / r
        dt = drt / r
        c = C.closure(method, beta, vr, t)
        c = fourier_rows(c) * dr ** 3 - (beta * A) * vk
        dc = fourier_rows(C.closure1(method, beta, vr, t, dt)) * dr ** 3
        ddt = oz_vv_equation_c_t1(rho, c, W, dc)
        ddt = fourier_rows(ddt) * (dk ** 3 / FT_BW)
        return r * (ddt - dt)

    jac = jacobian_rt if method != C.PY else None
    rt, stats = newton_krylov(iterate_rt, np.zeros((m, m, nrad)), J=jac,
                              atol=atol, verbose=verbose)
    t = rt / r
    c = C.closure(method, beta, vr, t)
    h = fourier_rows(c + t) * dr ** 3
    chi = wk + rho * h
    vl = force_field_long(sites, sites, r)
    mu = chempot(method, rmax, beta, rho, vr, vl, t)
    return {"t": t, "chi": chi, "mu": mu, "stats": stats}


def rism_uv(method, nrad, rmax, beta, rho, solvent, chi, solute, atol=1e-12,
            verbose=False):
    """rism.f90:586-859 (ng = .false. branch).  chi is (m, m, nrad)."""
    n, m = len(solute), len(solvent)
    r, dr, k, dk = mkgrid(rmax, nrad)
    v_uvr, v_uvk = force_field(solute, solvent, r, k)
    w_uuk = omega_fourier(solute, k)
    s_uvk = -beta * (oz_uv_equation_c_t(v_uvk, w_uuk, chi) + v_uvk)
    FT_BW = (2 * PI) ** 3

    def iterate_t(t):
        c = C.closure(method, beta, v_uvr, t)
        c = fourier_rows(c) * dr ** 3
        dt = oz_uv_equation_c_t(c, w_uuk, chi) + s_uvk
        dt = fourier_rows(dt) * (dk ** 3 / FT_BW)
        return dt - t

    def jacobian_t(t, dt):
        dc = C.closure1(method, beta, v_uvr, t, dt)
        dc = fourier_rows(dc) * dr ** 3
        ddt = oz_uv_equation_c_t(dc, w_uuk, chi)
        ddt = fourier_rows(ddt) * (dk ** 3 / FT_BW)
        return ddt - dt

    jac = jacobian_t if method != C.PY else None
    t, stats = newton_krylov(iterate_t, np.zeros((n, m, nrad)), J=jac,
                             atol=atol, verbose=verbose)
    vl = force_field_long(solute, solvent, r)
    mu = chempot(method, rmax, beta, rho, v_uvr, vl, t)
    return {"t": t, "mu": mu, "stats": stats}


def rism_solvent_renorm(solvent, rmax, nrad, chi, alpha=ALPHA):
    """rism.f90:141-208: s_i(k) = sum_j chi_ij(k) q_j 4pi e^{-k^2/4a^2}/k^2.
    chi (m, m, nrad) -> s (m, nrad)."""
    _r, _dr, k, _dk = mkgrid(rmax, nrad)
    q = np.array([s[4] for s in solvent], dtype=np.float64)
    v = coulomb_long_fourier(k, alpha)
    return np.einsum("ijk,j->ik", chi, q) * v[None, :]

def rism_vv_mixture(method, nrad, rmax, beta, sites, molecule, rho_sites, atol=1e-12,
                    verbose=False):
    """XRISM of a mixture (no counterpart in the reference, whose solvent is
    one molecule at one density, hnc3d.c:90-92): sites of several molecules,
    `molecule[i]` the molecule of site i, `rho_sites[i]` its number density.
    Per k, H = (1 - W C D)^-1 W C W with D = diag(rho), W block-diagonal by
    molecule; the closure and long-range handling as rism_vv (no DRISM).
    Returns dict(t, chi) with chi_ij = w_ij + rho_i h_ij, the matrix whose
    transpose contracts the solute's c: h_i = sum_j c_j chi_ji."""
    m = len(sites)
    rho = np.asarray(rho_sites, dtype=np.float64)
    if m == 1:
        rho = float(rho[0])
    r, dr, k, dk = mkgrid(rmax, nrad)
    vr, vk = force_field(sites, sites, r, k)
    wk = omega_fourier(sites, k, molecule)
    FT_BW = (2 * PI) ** 3

    def iterate_t(t):
        c = C.closure(method, beta, vr, t)
        c = fourier_rows(c) * dr ** 3 - beta * vk
        dt = oz_vv_equation_c_t(rho, c, wk) - beta * vk
        return fourier_rows(dt) * (dk ** 3 / FT_BW) - t

    def jacobian_rt(rt, drt):
        t = rt / r
        dt = drt / r
        c = fourier_rows(C.closure(method, beta, vr, t)) * dr ** 3 - beta * vk
        dc = fourier_rows(C.closure1(method, beta, vr, t, dt)) * dr ** 3
        ddt = fourier_rows(oz_vv_equation_c_t1(rho, c, wk, dc)) * (dk ** 3 / FT_BW)
        return r * (ddt - dt)

    rt, stats = newton_krylov(lambda x: r * iterate_t(x / r), np.zeros((m, m, nrad)),
                              J=jacobian_rt if method != C.PY else None, atol=atol,
                              verbose=verbose)
    t = rt / r
    c = C.closure(method, beta, vr, t)
    h = fourier_rows(c + t) * dr ** 3
    chi = wk + np.reshape(rho, (-1, 1, 1)) * h
    return {"t": t, "chi": chi, "stats": stats}


def renorm_mixture(solvent, rmax, nrad, chi, alpha=ALPHA):
    """tau_i(k) = sum_j q_j chi_ji(k) 4pi e^{-k^2/4a^2}/k^2: the long-range
    response of site i to the solute field through every site j, the
    transposed contraction of rism_solvent_renorm (the same for a symmetric
    chi)."""
    _r, _dr, k, _dk = mkgrid(rmax, nrad)
    q = np.array([s[4] for s in solvent], dtype=np.float64)
    return np.einsum("jik,j->ik", chi, q) * coulomb_long_fourier(k, alpha)[None, :]
