"""Gradients of the solvation free energy (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

* ``tpt_gradient``  restates the reference's --derivatives path:
  gradients() -> energy1() (hnc3d.c:1666-1742) with field1()
  (bgy3d-solutes.c:381-468): de/dx_a = rho h^3 sum_i <g_i | dv_i/dx_a>,
  where dv uses the capped LJ derivative ljcap1 and the capped *full*
  Coulomb derivative ccap1 (bgy3d-solutes.c:102-131, 290-305).
  Pinned: dv/dx against central differences of the full site potential.

* ``response_gradient`` restates the reference's --response path
  (hnc3d.c:1366-1421 response_t1, hnc3d.c:1969-2134): for every solute
  site and axis, dv = field1 (full capped Coulomb), b = beta J' dv with
  the Jacobian's flag hack (hnc3d.c:1310-1320), J dt = b solved to a
  relative tolerance, then chempot1 (closures.f90:615-653) of dx = -beta dv
  + dt, dc = c1(dx) - beta dv, dh = dc + dt at fixed cl.  3 n linear
  solves.  Pinned: equal to central differences of mu for an uncharged
  solute (the long-range terms it leaves out vanish there), and to the
  one-solve adjoint form sum_i <A_i | dv_i> for charged ones.

* ``adjoint_gradient`` is exact d mu / d(x, y, z, q, sigma, epsilon) of
  every solute site on the renormalised and the Ng path, by the adjoint of
  the residual: with J = dF/dt = K D_t - 1 (K the chi convolution,
  symmetric because chi(|k|) is real and even) the multiplier solves
  lambda = D_t K lambda + dmu/dt, and
  dmu/dtheta = dmu/dtheta|_t + lambda . dF/dtheta|_t.  The long-range term
  of F is -(beta eps0^-1 / L^3) w_i(k) F(k) s(k) on both paths: w_i = tau_i
  (renorm) or 4 pi e^{-k^2/4G^2}/k^2 (q_i + sum_j (chi - 1)_ij q_j), 0 at
  k = 0 (Ng, hnc3d.c:1226-1227, 1258-1262).
  There is no reference counterpart (the reference has position
  derivatives only, as TPT or by linear response); pinned by central
  finite differences of converged mu in tests/test_oracle_gradients.py.
"""
import math

import numpy as np
from scipy.special import erf, erfc

from . import closures as C
from . import rism3d as O
from .nksolve import mdiis

EPS0INV = O.EPSILON0INV
G = O.G
RMIN = 0.2


def ljcap1(r, R=RMIN):
    b = O.lj1(R) / (2 * R)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        f = O.lj1(r)
    return np.where(r >= R, f, 2 * b * r)


def cs1(r):
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        return -(erfc(r) + (2 / math.sqrt(math.pi)) * r * np.exp(-r * r)) / (r * r)


def cscap1(r, R=RMIN):
    b = O.cs1(R) / (2 * R)
    return np.where(r >= R, cs1(r), 2 * b * r)


def cl1(r):
    r2 = r * r
    small = -(4 / (3 * math.sqrt(math.pi))) * r * (1. + r2 * (-3. / 5. + r2 * (3. / 14. + r2 * (-1. / 18.))))
    with np.errstate(divide="ignore", invalid="ignore"):
        big = ((2 / math.sqrt(math.pi)) * r * np.exp(-r2) - erf(r)) / r2
    return np.where(r >= 0.02, big, small)


def ccap0(r, R=RMIN):
    with np.errstate(divide="ignore"):
        return np.where(r >= R, 1.0 / r, O.cscap0(r) + O.cl0(r))


def ccap1(r, R=RMIN):
    with np.errstate(divide="ignore"):
        return np.where(r >= R, -1.0 / (r * r), cscap1(r) + cl1(r))


def _rvec(pd, xyz):
    x, y, z = O.r_axes(pd)
    return ((x - xyz[0])[None, None, :], (y - xyz[1])[None, :, None], (z - xyz[2])[:, None, None])


def full_field(pd, a, solute):
    """Capped LJ + capped full Coulomb potential of solvent site a, the
    function field1 differentiates."""
    v = np.zeros(pd.N[::-1])
    for b in solute:
        eab = math.sqrt(a[3] * b[3])
        sab = 0.5 * (a[2] + b[2])
        qab = a[4] * b[4] * EPS0INV
        r = O._site_dist(pd, b[1])
        v += eab * O.ljcap0(r / sab) + qab * G * ccap0(G * r)
    return v


def field1(pd, a, b, d):
    """d v_a / d x_b[d] of solvent site a in the field of solute site b
    (bgy3d-solutes.c:381-468 with one moving site): capped LJ and capped
    full Coulomb derivatives times d r / d x_b."""
    dx, dy, dz = _rvec(pd, b[1])
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(r > 0, 1.0 / r, 0.0)
    eab = math.sqrt(a[3] * b[3])
    sab = 0.5 * (a[2] + b[2])
    qab = a[4] * b[4] * EPS0INV
    f = (eab / sab) * ljcap1(r / sab) + qab * G * G * ccap1(G * r)
    return f * (-(dx, dy, dz)[d] * inv)


def tpt_gradient(pd, solvent, solute, g):
    """gradients() of hnc3d.c:1726-1742: de[n][3]."""
    dN = pd.rho * pd.h3
    de = np.zeros((len(solute), 3))
    for a, b in enumerate(solute):
        dx, dy, dz = _rvec(pd, b[1])
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.where(r > 0, 1.0 / r, 0.0)
        for i, s in enumerate(solvent):
            eab = math.sqrt(s[3] * b[3])
            sab = 0.5 * (s[2] + b[2])
            qab = s[4] * b[4] * EPS0INV
            f = (eab / sab) * ljcap1(r / sab) + qab * G * G * ccap1(G * r)
            for d, comp in enumerate((dx, dy, dz)):
                de[a, d] += dN * np.sum(g[i] * f * (-comp * inv))
    return de


def _mu_fields(sol, t):
    pd = sol.pd
    beta = pd.beta
    c = C.closure(pd.closure, beta, sol.v, t)
    h = c + t
    x = -beta * sol.v + t
    uc = O.coulomb_long(pd, sol.solute)
    cl = -beta * sol.q[:, None, None, None] * uc[None]
    return x, h, c, cl


def response_gradient(sol, t, rel_tol=1e-12):
    """'free-energy-response of hnc3d.c:1969-2134: dE[n][3]."""
    pd = sol.pd
    beta = pd.beta
    if pd.closure == C.PY:
        raise ValueError("the reference's closure1 has no PY branch")
    fmu = pd.rho * pd.h3 / beta
    x, h, c, cl = _mu_fields(sol, t)
    dE = np.zeros((len(sol.solute), 3))
    for a, b in enumerate(sol.solute):
        for d in range(3):
            dv = np.stack([field1(pd, s, b, d) for s in sol.solvent])
            rhs = beta * sol.jacobian(t, dv, flag=True)
            atol = rel_tol * max(np.max(np.abs(rhs)), 1e-300)
            dt, _ = mdiis(lambda y: sol.jacobian(t, y) - rhs, np.zeros_like(t),
                          atol=atol, max_iter=5000)
            dx = -beta * dv + dt
            dc = C.closure1(pd.closure, beta, sol.v, t, dx) - beta * dv
            dh = dc + dt
            dE[a, d] = fmu * math.fsum(C.form1(pd.closure, x, dx, h, dh, c, dc, cl).ravel())
    return dE


def response_from_adjoint(sol, t, lam):
    """The same numbers from one adjoint solve: with lambda = D_t K lambda +
    dmu/dt, dE[a][d] = sum_i <A_i | field1_i>, A_i = -beta (ct + 1) K lambda
    + dmu/dv at fixed t (ct = dc/dt)."""
    pd = sol.pd
    beta = pd.beta
    fmu = pd.rho * pd.h3 / beta
    x, h, c, cl = _mu_fields(sol, t)
    ct = C.closure1(pd.closure, beta, sol.v, t, np.ones_like(t))
    Klam = O.irfft(pd, sol.star(O.rfft(pd, lam) * pd.h3)) / pd.L3
    dcv = -beta * (ct + 1)
    A = Klam * dcv + fmu * C.form1(pd.closure, x, -beta, h, dcv, c, dcv, cl)
    dE = np.zeros((len(sol.solute), 3))
    for a, b in enumerate(sol.solute):
        for d in range(3):
            dE[a, d] = sum(np.sum(A[i] * field1(pd, s, b, d)) for i, s in enumerate(sol.solvent))
    return dE


def adjoint_lambda(sol, t, rel_tol=1e-12):
    """lambda = D_t K lambda + dmu/dt (any closure but PY, either path)."""
    pd = sol.pd
    fmu = pd.rho * pd.h3 / pd.beta
    x, h, c, cl = _mu_fields(sol, t)
    one = np.ones_like(t)
    ct = C.closure1(pd.closure, pd.beta, sol.v, t, one)
    gmu = fmu * C.form1(pd.closure, x, one, h, ct + 1, c, ct, cl)

    def K(v):
        return O.irfft(pd, sol.star(O.rfft(pd, v) * pd.h3)) / pd.L3

    lam, _ = mdiis(lambda l: ct * K(l) + gmu - l, np.zeros_like(t),
                   atol=rel_tol * np.max(np.abs(gmu)), max_iter=5000)
    return lam


def _kweights(pd):
    w = np.full(pd.N[0] // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return w[None, None, :]


def adjoint_gradient(sol, t, lam_rel_tol=1e-12):
    """Exact dmu/dtheta for a converged Solute3D ``sol`` (either path) at t.
    Returns (grad[n][6] over x, y, z, q, sigma, epsilon; lambda)."""
    pd = sol.pd
    beta, rho = pd.beta, pd.rho
    meth = pd.closure
    fmu = rho * pd.h3 / beta
    bv = beta * sol.v
    c = C.closure(meth, beta, sol.v, t)
    h = c + t
    x = -bv + t
    uc = O.coulomb_long(pd, sol.solute)
    q = sol.q[:, None, None, None]
    cl = -beta * q * uc[None]
    one = np.ones_like(t)
    ct = C.closure1(meth, beta, sol.v, t, one)          # dc/dt
    # dc/d(beta v): c = f(-bv + t) - t except KH x > 0 (c = -bv) and PY
    if meth == C.PY:
        cv = -np.exp(-bv) * (1 + t)
    elif meth == C.KH:
        cv = np.where(x <= 0, -(ct + 1), -1.0)
    else:
        cv = -(ct + 1)
    gmu = fmu * C.form1(meth, x, one, h, ct + 1, c, ct, cl)

    def K(v):
        return O.irfft(pd, sol.star(O.rfft(pd, v) * pd.h3)) / pd.L3

    lam, _ = mdiis(lambda l: ct * K(l) + gmu - l, np.zeros_like(t),
                   atol=lam_rel_tol * np.max(np.abs(gmu)), max_iter=5000)
    Klam = K(lam)
    # coefficient of dv_i(r): lambda . K dc + dM/dv at fixed t
    dx_ = -beta * one
    dc_ = cv * beta
    A = beta * Klam * cv + fmu * C.form1(meth, x, dx_, h, dc_, c, dc_, cl)
    B = 0.5 * rho * pd.h3 * np.sum(q * h, axis=0)        # dM/d u_c(r)
    # tau term: -tau_scale Re sum_k w tau_i s dF/dtheta conj(Lambda_i)
    Lam = O.rfft(pd, lam)
    s = O.phase_sign(pd)
    if sol.renorm:
        tau0 = sol.tau_k
    else:
        k = O.kmag(pd)
        with np.errstate(divide="ignore"):
            u0 = 4 * math.pi * np.exp(-k * k / (4 * G * G)) / (k * k)
        u0 = np.where(k == 0.0, 0.0, u0)
        tau0 = u0[None] * (sol.q[:, None, None, None] + np.einsum("ij...,j->i...", sol.chi, sol.q))
    S = np.sum(tau0 * np.conj(Lam), axis=0) * s * _kweights(pd)
    bscale = beta * EPS0INV / pd.L3
    kx, ky, kz = O.k_axes(pd)
    n = len(sol.solute)
    grad = np.zeros((n, 6))
    for a, b in enumerate(sol.solute):
        dx, dy, dz = _rvec(pd, b[1])
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.where(r > 0, 1.0 / r, 0.0)
        drdx = [-dx * inv, -dy * inv, -dz * inv]
        Qa, sa, ea = b[4], b[2], b[3]
        for i, sv in enumerate(sol.solvent):
            eab = math.sqrt(sv[3] * ea)
            sab = 0.5 * (sv[2] + sa)
            qi = sv[4]
            rs = r / sab
            fr = (eab / sab) * ljcap1(rs) + qi * Qa * EPS0INV * G * G * cscap1(G * r)
            for d in range(3):
                grad[a, d] += np.sum(A[i] * fr * drdx[d])
            grad[a, 3] += np.sum(A[i] * qi * EPS0INV * G * O.cscap0(G * r))
            grad[a, 4] += np.sum(A[i] * eab * ljcap1(rs) * (-r / sab ** 2) * 0.5)
            if eab > 0:
                grad[a, 5] += np.sum(A[i] * O.ljcap0(rs) * sv[3] / (2 * eab))
        # u_c(r) in mu: d uc/d theta
        for d in range(3):
            grad[a, d] += np.sum(B * EPS0INV * Qa * G * G * cl1(G * r) * drdx[d])
        grad[a, 3] += np.sum(B * EPS0INV * G * O.cl0(G * r))
        # tau term through F(k) = sum Q e^{-ik.x}
        ph = (np.exp(-1j * kz * b[1][2])[:, None, None] * np.exp(-1j * ky * b[1][1])[None, :, None]
              * np.exp(-1j * kx * b[1][0])[None, None, :])
        grad[a, 3] += -bscale * np.real(np.sum(S * ph))
        for d, kd in enumerate((kx[None, None, :], ky[None, :, None], kz[:, None, None])):
            grad[a, d] += -bscale * np.real(np.sum(S * ph * (-1j) * kd * Qa))
    return grad, lam
