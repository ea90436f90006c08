"""Newton-GMRES with backtracking line search (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Stands in for PETSc SNESNEWTONLS + KSPGMRES(restart 30, PCNONE) as set up in
bgy3d-snes.c:161-346 (bgy3d_snes_newton).  PETSc's exact iteration path is
not reproducible here (PETSc is absent), and it does not need to be: every
solve in this path is a root-finding problem F(x) = 0 whose root is unique
for the cases tested, so only the tolerance matters for parity.

Where no Jacobian is given (1D solvent-solvent RISM, rism.f90:413-420, which
PETSc handles with MatMFFD), J.v is a forward finite difference.
"""
import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres


class NotConverged(RuntimeError):
    pass


def mdiis(F, x0, atol, depth=10, lam=0.5, max_iter=3000):
    """Pulay/MDIIS fixed-point iteration x <- sum c_j (x_j + lam F(x_j)),
    used to bring Newton into its basin on coarse grids.  Returns (x, it)."""
    xs, rs = [], []
    x = np.array(x0, dtype=np.float64, copy=True)
    best = np.inf
    for it in range(max_iter):
        r = F(x)
        rn = np.max(np.abs(r))
        if not np.isfinite(rn):
            raise NotConverged("MDIIS residual is not finite")
        if rn <= atol:
            return x, it
        r2 = float(np.vdot(r, r))
        if r2 > 100 * best:
            xs, rs = [], []
        best = min(best, r2)
        xs.append(x)
        rs.append(r)
        if len(xs) > depth:
            xs.pop(0)
            rs.pop(0)
        k = len(rs)
        A = np.zeros((k + 1, k + 1))
        A[:k, :k] = [[float(np.vdot(a, b)) for b in rs] for a in rs]
        A[:k, :k] /= np.max(np.abs(np.diag(A[:k, :k])))
        A[k, :k] = A[:k, k] = -1.0
        rhs = np.zeros(k + 1)
        rhs[k] = -1.0
        try:
            c = np.linalg.solve(A, rhs)[:k]
        except np.linalg.LinAlgError:
            xs, rs, c = [xs[-1]], [rs[-1]], np.array([1.0])
        x = sum(ci * (xi + lam * ri) for ci, xi, ri in zip(c, xs, rs))
    raise NotConverged(f"MDIIS: no convergence after {max_iter} iterations")


def newton_krylov(F, x0, J=None, atol=1e-12, max_iter=100, ksp_rtol=1e-4,
                  restart=30, ksp_maxiter=300, verbose=False):
    """Solve F(x) = 0 from x0.  Stops when max|F| <= atol.

    F: x -> residual (same shape).  J: (x, dx) -> J(x) dx or None.
    Returns (x, stats) with stats = dict(n_f, n_j, iters, resid)."""
    x = np.array(x0, dtype=np.float64, copy=True)
    shape = x.shape
    n = x.size
    stats = {"n_f": 0, "n_j": 0, "iters": 0}

    def f(z):
        stats["n_f"] += 1
        return F(z.reshape(shape)).reshape(-1)

    z = x.reshape(-1)
    r = f(z)
    for it in range(max_iter):
        rn = np.max(np.abs(r))
        if verbose:
            print(f"  newton {it:3d} |F|inf = {rn:.3e}")
        if not np.isfinite(rn):
            raise NotConverged("residual is not finite")
        if rn <= atol:
            stats["iters"] = it
            stats["resid"] = rn
            return z.reshape(shape), stats
        if J is not None:
            def mv(v, z=z):
                stats["n_j"] += 1
                return J(z.reshape(shape), v.reshape(shape)).reshape(-1)
        else:
            def mv(v, z=z, r=r):
                stats["n_j"] += 1
                nv = np.linalg.norm(v)
                if nv == 0.0:
                    return np.zeros_like(v)
                eps = 1e-7 * (1.0 + np.linalg.norm(z)) / nv
                return (f(z + eps * v) - r) / eps
        A = LinearOperator((n, n), matvec=mv, dtype=np.float64)
        dz, _info = gmres(A, -r, rtol=ksp_rtol, atol=0.0, restart=restart,
                          maxiter=ksp_maxiter)
        # Backtracking line search on 0.5|F|^2 (PETSc "bt" flavour):
        lam = 1.0
        f0 = 0.5 * np.dot(r, r)
        while True:
            zn = z + lam * dz
            rn_vec = f(zn)
            fn = 0.5 * np.dot(rn_vec, rn_vec)
            if np.isfinite(fn) and fn <= (1.0 - 1e-4 * lam) * f0:
                break
            lam *= 0.5
            if lam < 1e-6:
                break
        z, r = zn, rn_vec
    raise NotConverged(f"no convergence after {max_iter} Newton steps, "
                       f"|F|inf = {np.max(np.abs(r)):.3e}")
