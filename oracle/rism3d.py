"""3D-RISM solute-in-solvent solve (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates hnc3d_solute_solve (hnc3d.c:1762-2350) and everything it calls on
the renormalised (default) and --no-renorm paths:

* grid, r_p = p h - L/2             bgy3d-vec.h:579-604 (vec_rmap3)
* k grid, KFREQ corner origin       bgy3d-vec.h:611-645 (vec_kmap3), bgy3d.h:162
* s(k) = (-1)^(kx+ky+kz)            bgy3d-vec.c:619-646 (bgy3d_vec_fft_trans)
* interp / vec_ktab                 bgy3d-vec.c:23-68
* field0 (LJ cap + erfc cap)        bgy3d-solutes.c:344-377, ljcap0 102,
                                    cscap0 231, cl0 148
* coulomb_long u_c(r)               bgy3d-solutes.c:796-815
* coulomb_long_fft u_c(k)           bgy3d-solutes.c:767-792, bgy3d-force.c:19-38
* form factor / solute_form         bgy3d-solutes.c:587-655
* solvent_kernel_rism (chi-1, tau)  hnc3d.c:1443-1566
* iterate_t1 / jacobian_t1          hnc3d.c:1184-1362
* chempot / pmv_factor              hnc3d.c:486-623, 702-737, 2172-2268

Arrays follow the reference DMDA layout: real [N2][N1][N0] (x fastest),
complex [N2][N1][N0/2+1].  FFTW r2c (sign -1, unnormalised) is
numpy.fft.rfftn over axes (0,1,2); FFTW c2r (unnormalised) is
irfftn(norm="forward"), which shares FFTW's treatment of the kx = 0 and
kx = N0/2 planes (complex inverse over z, y, then real part at DC/Nyquist).

Pinning: README:20-31 (-6.349052 renorm, -6.297170 --no-renorm for
PR-SPC/E water at --L 10 --N 96) reproduced to 1e-7 with chi from 1D DRISM
on the grid hnc3d_solute_solve builds itself when chi_fft_buf is NULL
(upscale(): rmax = 40, nrad = 1536, hnc3d.c:772-783);
test/out/run-ions-out:56-110 (nine charged single sites, KH/PSE-n/HNC):
per closure every ion to the printed 4 decimals after one constant of the
reference's own solvent chi run, which the 1D rows show as well (13-15
times smaller); a dropped tau(0) F(0) term fails that by 0.14 kcal/mol;
grid invariants; in tests/test_oracle_*.py.
"""
import math
import os

import numpy as np
import scipy.fft
from scipy.special import erf, erfc

from . import closures as C
from . import rism1d
from .nksolve import NotConverged, mdiis, newton_krylov

EPSILON0INV = 331.84164
G = 1.2


class Problem:
    """ProblemData (bgy3d.h:228-256), derived as problem_data()
    (bgy3d-guile.c:243-385): L full edge = 2*length, h = L/N,
    rmax = max(L)/2, nrad = max(N) unless given."""

    def __init__(self, N, L, beta, rho, closure=C.HNC, rmax=None, nrad=None):
        self.N = (N, N, N) if np.isscalar(N) else tuple(N)
        self.L = (float(L),) * 3 if np.isscalar(L) else tuple(map(float, L))
        self.h = tuple(l / n for l, n in zip(self.L, self.N))
        self.beta = float(beta)
        self.rho = float(rho)
        self.closure = closure
        self.rmax = max(self.L) / 2 if rmax is None else float(rmax)
        self.nrad = max(self.N) if nrad is None else int(nrad)

    @classmethod
    def from_cli(cls, length, N, beta, rho, closure=C.HNC, rmax=None, nrad=None):
        """--L is the half edge (bgy3d-guile.c:284-288)."""
        return cls(N, 2.0 * length, beta, rho, closure, rmax, nrad)

    @property
    def h3(self):
        return self.h[0] * self.h[1] * self.h[2]

    @property
    def L3(self):
        return self.L[0] * self.L[1] * self.L[2]


def kfreq(n):
    i = np.arange(n)
    return np.where(i <= n // 2, i, i - n)


def r_axes(pd):
    return [np.arange(n) * h - l / 2 for n, h, l in zip(pd.N, pd.h, pd.L)]


def k_axes(pd):
    """kx over 0..N0/2 (half-complex), ky, kz full."""
    N0, N1, N2 = pd.N
    kx = np.arange(N0 // 2 + 1) * (2 * math.pi / pd.L[0])
    ky = kfreq(N1) * (2 * math.pi / pd.L[1])
    kz = kfreq(N2) * (2 * math.pi / pd.L[2])
    return kx, ky, kz


def kmag(pd):
    """|k| = sqrt((kx^2 + ky^2) + kz^2) in the reference's order
    (SQR(k[0]) + SQR(k[1]) + SQR(k[2]), bgy3d-vec.h:664-667).  The order
    matters: interp (bgy3d-vec.c:23-46) is discontinuous at table midpoints
    and grid k often lands exactly on one (|k|/dk_tab integer), so the
    last bit of |k|^2 picks the neighbour."""
    kx, ky, kz = k_axes(pd)
    k2 = (kx[None, None, :] * kx[None, None, :] + ky[None, :, None] * ky[None, :, None]) \
        + kz[:, None, None] * kz[:, None, None]
    return np.sqrt(k2)


def phase_sign(pd):
    N0, N1, N2 = pd.N
    sx = np.where(np.arange(N0 // 2 + 1) % 2, -1.0, 1.0)
    sy = np.where(kfreq(N1) % 2, -1.0, 1.0)
    sz = np.where(kfreq(N2) % 2, -1.0, 1.0)
    return sz[:, None, None] * sy[None, :, None] * sx[None, None, :]


def interp(x, tab, dx):
    """bgy3d-vec.c:23-46, vectorised."""
    n = tab.size
    ix = x / dx - 0.5
    i = np.floor(ix + 0.5).astype(np.int64)     # (int)(ix + 0.5), ix+0.5 >= 0
    if np.any(i < 0) or np.any(i >= n):
        raise ValueError("k table too short for this grid (interp asserts)")
    ip = np.where(i < n - 2, i + 1, i)
    im = np.where(i > 1, i - 1, i)
    fi = (tab[ip] - tab[im]) / (ip - im)
    return tab[i] + fi * (ix - i)


def ktab(pd, table, dk):
    return interp(kmag(pd), np.asarray(table, dtype=np.float64), dk)


# ---- solute fields (bgy3d-solutes.c) ---------------------------------

def lj0(r):
    p6 = r ** -6.0
    return 4 * p6 * (p6 - 1)


def lj1(r):
    p6 = r ** -6.0
    return 4 * p6 * (6 - 12 * p6) / r


def ljcap0(r, R=0.2):
    b = lj1(R) / (2 * R)
    a = lj0(R)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        f = lj0(r)
    return np.where(r >= R, f, a + b * (r - R) * (r + R))


def cs0(r):
    return erfc(r) / r


def cs1(r):
    return -(erfc(r) + (2 / math.sqrt(math.pi)) * r * math.exp(-r * r)) / (r * r)


def cscap0(r, R=0.2):
    b = cs1(R) / (2 * R)
    a = cs0(R)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = erfc(r) / r
    return np.where(r >= R, f, a + b * (r - R) * (r + R))


def cl0(r):
    r2 = r * r
    small = (2 / math.sqrt(math.pi)) * (1. + r2 * (-1. / 3. + r2 * (1. / 10. + r2 * (-1. / 42.))))
    with np.errstate(divide="ignore", invalid="ignore"):
        big = erf(r) / r
    return np.where(r >= 0.02, big, small)


def _site_dist(pd, xyz):
    x, y, z = r_axes(pd)
    dz = (z - xyz[2])[:, None, None]
    dy = (y - xyz[1])[None, :, None]
    dx = (x - xyz[0])[None, None, :]
    return np.sqrt(dx * dx + dy * dy + dz * dz)


def field0(pd, a, solute):
    """Short-range field acting on solvent site a (bgy3d-solutes.c:344-377)."""
    v = np.zeros(pd.N[::-1])
    for b in solute:
        eab = math.sqrt(a[3] * b[3])
        sab = 0.5 * (a[2] + b[2])
        qab = a[4] * b[4] * EPSILON0INV
        rab = _site_dist(pd, b[1])
        v += eab * ljcap0(rab / sab) + qab * G * cscap0(G * rab)
    return v


def coulomb_long(pd, solute):
    u = np.zeros(pd.N[::-1])
    for b in solute:
        u += b[4] * G * cl0(G * _site_dist(pd, b[1]))
    return EPSILON0INV * u


def form_factor(pd, solute):
    """F(k) = sum_a Q_a exp(-i k.x_a), corner k (bgy3d-solutes.c:587-600)."""
    kx, ky, kz = k_axes(pd)
    f = np.zeros((pd.N[2], pd.N[1], pd.N[0] // 2 + 1), dtype=np.complex128)
    for b in solute:
        x = b[1]
        f += b[4] * (np.exp(-1j * kz * x[2])[:, None, None]
                     * np.exp(-1j * ky * x[1])[None, :, None]
                     * np.exp(-1j * kx * x[0])[None, None, :])
    return f


def coulomb_long_fft(pd, solute):
    """bgy3d-solutes.c:767-792: unit-shape centred, times centred form
    factor, times s(k) once more -> net F(k) * shape(k) (corner)."""
    k = kmag(pd)
    with np.errstate(divide="ignore"):
        shape = (4 * math.pi * EPSILON0INV) * np.exp(-k * k / (4 * G * G)) / (k * k)
    shape = np.where(k == 0.0, 0.0, shape)
    s = phase_sign(pd)
    return shape * s * (form_factor(pd, solute) * s) * s


def solvent_kernel_rism(pd, solvent, chi_1d):
    """hnc3d.c:1443-1566 with a caller-supplied chi (m, m, nrad)."""
    m = len(solvent)
    dk = math.pi / pd.rmax
    s_kv = rism1d.rism_solvent_renorm(solvent, pd.rmax, pd.nrad, chi_1d)
    tau = np.stack([ktab(pd, s_kv[i], dk) for i in range(m)]).astype(np.complex128)
    chi = np.zeros((m, m) + tau.shape[1:])
    for i in range(m):
        for j in range(i + 1):
            tab = chi_1d[i, j] - (1.0 if i == j else 0.0)
            chi[i, j] = ktab(pd, tab, dk)
            chi[j, i] = chi[i, j]
    return chi, tau


# pocketfft through scipy.fft on up to 16 host threads (the same transforms
# as numpy.fft, which runs them on one thread)
FFT_WORKERS = max(1, min(16, os.cpu_count() or 1))
# The closure and the chi contraction split over slabs of z planes on this
# many threads (numpy releases the GIL inside its loops); 1 runs them on
# the calling thread.  Same arithmetic either way, point by point.
ELEMENTWISE_WORKERS = 1
_POOL = None


def _zsplit(fn, zaxes, *arrays):
    """fn(*slabs) over slabs of z planes, joined along the output's z axis
    (axis 1 of an [m][N2][N1][N0] result); zaxes[i] is the z axis of
    arrays[i]."""
    global _POOL
    nz = arrays[0].shape[zaxes[0]]
    nw = min(ELEMENTWISE_WORKERS, nz)
    if nw <= 1:
        return fn(*arrays)
    if _POOL is None or _POOL._max_workers < nw:
        from concurrent.futures import ThreadPoolExecutor
        _POOL = ThreadPoolExecutor(max_workers=nw)
    cuts = np.linspace(0, nz, nw + 1).astype(int)

    def part(k):
        z = slice(cuts[k], cuts[k + 1])
        return fn(*[a[(slice(None),) * ax + (z,)] for a, ax in zip(arrays, zaxes)])

    return np.concatenate(list(_POOL.map(part, range(nw))), axis=1)


def rfft(pd, a):
    return scipy.fft.rfftn(a, axes=(-3, -2, -1), workers=FFT_WORKERS)


def irfft(pd, a):
    """Unnormalised c2r (FFTW semantics)."""
    return scipy.fft.irfftn(a, s=pd.N[::-1], axes=(-3, -2, -1), norm="forward",
                            workers=FFT_WORKERS)


class Solute3D:
    """Everything hnc3d_solute_solve sets up before SNESSolve."""

    def __init__(self, pd, solvent, solute, chi_1d, renorm=True, chi3d=None, site_rho=None):
        """chi3d: None, or chi - 1 of a 3D solvent solve on this grid,
        [m(m+1)/2][N2][N1][N0/2+1] by pair i(i+1)/2 + j (j <= i): the
        --solvent-3d path (solvent_kernel_file, hnc3d.c:1434-1437 and
        1572-1590), which has no tau renormalisation.
        site_rho: None (every site at pd.rho, as the reference), or the
        number density of each solvent site of a mixture (no counterpart in
        the reference): chi_1d is then the non-symmetric
        chi_ij = w_ij + rho_i h_ij of rism1d.rism_vv_mixture, the solute's
        c is contracted as h_i = sum_j c_j chi_ji, tau_i = sum_j q_j chi_ji
        (rism1d.renorm_mixture), and mu and the PMV factor weight site i
        by rho_i."""
        self.pd = pd
        self.solvent = solvent
        self.solute = solute
        self.m = len(solvent)
        self.q = np.array([s[4] for s in solvent])
        self.rho_sites = (np.full(self.m, pd.rho) if site_rho is None
                          else np.asarray(site_rho, dtype=np.float64))
        if site_rho is not None:
            m = self.m
            dk = math.pi / pd.rmax
            s_kv = rism1d.renorm_mixture(solvent, pd.rmax, pd.nrad, chi_1d)
            tau = np.stack([ktab(pd, s_kv[i], dk) for i in range(m)]).astype(np.complex128)
            # chi[i, j] holds chi_ji - delta_ij: star() then contracts as the
            # transposed matrix
            self.chi = np.stack([np.stack([ktab(pd, chi_1d[j, i] - (1.0 if i == j else 0.0), dk)
                                           for j in range(m)]) for i in range(m)])
        elif chi3d is not None:
            m = self.m
            self.chi = np.empty((m, m) + chi3d.shape[1:])
            for i in range(m):
                for j in range(i + 1):
                    self.chi[i, j] = self.chi[j, i] = chi3d[i * (i + 1) // 2 + j]
            tau = None
            renorm = False
        else:
            self.chi, tau = solvent_kernel_rism(pd, solvent, chi_1d)
        self.tau_k = tau                    # tau_i(|k|) before F(k) s(k)
        self.renorm = renorm
        if renorm:
            s = phase_sign(pd)
            self.tau = tau * (form_factor(pd, solute) * s)[None]
            self.uc_fft = None
        else:
            self.tau = None
            self.uc_fft = coulomb_long_fft(pd, solute)
        self.v = np.stack([field0(pd, a, solute) for a in solvent])

    def _closure(self, t):
        return _zsplit(lambda v, tt: C.closure(self.pd.closure, self.pd.beta, v, tt), (1, 1),
                       self.v, t)

    def star(self, x):
        return _zsplit(lambda chi, xx: np.einsum("ijzyx,jzyx->izyx", chi, xx), (2, 1),
                       self.chi, x)

    def iterate(self, t):
        """F(t) = T(t) - t (hnc3d.c:1184-1285)."""
        pd = self.pd
        beta = pd.beta
        c = self._closure(t)
        cf = rfft(pd, c) * pd.h3
        if not self.renorm:
            cf = cf - beta * self.q[:, None, None, None] * self.uc_fft[None]
        tf = self.star(cf)
        if self.renorm:
            tf = tf - beta * EPSILON0INV * self.tau
        else:
            tf = tf - beta * self.q[:, None, None, None] * self.uc_fft[None]
        return irfft(pd, tf) / pd.L3 - t

    def jacobian(self, t, dt, flag=False):
        """J dt (hnc3d.c:1290-1362)."""
        pd = self.pd
        dc = C.closure1(pd.closure, pd.beta, self.v, t, dt)
        if flag:
            dc = dc + dt
        cf = rfft(pd, dc) * pd.h3
        out = irfft(pd, self.star(cf)) / pd.L3
        return out if flag else out - dt

    def ladder(self):
        """Closures tried before the target when its iteration diverges:
        KH (PSE-1), PSE-2, ... (PSE-4 last before HNC).  Not in the
        reference; only the starting point of the target's iteration
        changes, not its root."""
        cl = self.pd.closure
        if cl == C.HNC:
            top = 5
        elif C.PSE0 + 2 <= cl <= C.PSE7:
            top = cl - C.PSE0
        else:
            return []
        return [C.KH] + [C.PSE0 + n for n in range(2, top)]

    def solve(self, t0=None, atol=1e-11, verbose=False, continuation=True):
        """MDIIS to atol; if it stalls, Newton-GMRES from its best iterate.
        If it diverges (non-finite residual), the closure ladder from t = 0,
        each rung to max|F| <= 1e-2.  Only the root matters for parity (see
        nksolve)."""
        t0 = np.zeros_like(self.v) if t0 is None else t0
        try:
            t, n_pre = mdiis(self.iterate, t0, atol=atol, max_iter=1500)
            return t, {"n_pre": n_pre, "iters": 0, "n_f": n_pre, "n_j": 0}
        except NotConverged as e:
            rungs = self.ladder() if continuation else []
            if rungs and "not finite" in str(e):
                target = self.pd.closure
                t0 = np.zeros_like(self.v)
                try:
                    for cl in rungs:
                        self.pd.closure = cl
                        t0, _ = mdiis(self.iterate, t0, atol=1e-2, max_iter=1500)
                finally:
                    self.pd.closure = target
                return self.solve(t0=t0, atol=atol, verbose=verbose, continuation=False)
            t0, n_pre = mdiis(self.iterate, t0, atol=1e-4)
        jac = self.jacobian if self.pd.closure != C.PY else None
        t, stats = newton_krylov(self.iterate, t0, J=jac, atol=atol,
                                 verbose=verbose, ksp_rtol=1e-3)
        stats["n_pre"] = n_pre
        return t, stats

    def observables(self, t):
        """mu (native closure), the HNC/KH/GF printout, PMV factor, g
        (hnc3d.c:2157-2321)."""
        pd = self.pd
        beta = pd.beta
        c = self._closure(t)
        h = c + t
        uc = coulomb_long(pd, self.solute)
        cl = -beta * self.q[:, None, None, None] * uc[None]
        x = -beta * self.v + t

        rs = self.rho_sites

        def mu_of(method):
            dens = C.form(method, x, h, c, cl)
            return math.fsum(rs[i] * math.fsum(dens[i].ravel()) for i in range(self.m)) * pd.h3 / beta

        kc = sum(rs[i] * math.fsum(c[i].ravel()) for i in range(self.m)) * pd.h3
        return {
            "mu": mu_of(pd.closure),
            "mu_hnc": mu_of(C.HNC), "mu_kh": mu_of(C.KH), "mu_gf": mu_of(C.PY),
            "pmv_factor": 1.0 - kc,
            "g": h + 1.0,
        }


def solute_solve(pd, solvent, solute, chi_1d, renorm=True, atol=1e-11,
                 verbose=False, t0=None):
    s = Solute3D(pd, solvent, solute, chi_1d, renorm=renorm)
    t, stats = s.solve(t0=t0, atol=atol, verbose=verbose)
    out = s.observables(t)
    out["t"] = t
    out["stats"] = stats
    return out
