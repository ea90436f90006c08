"""Python host layer over ``librism3d.so`` (the C ABI of include/rism3d.h).

Mirrors the reference's Guile-facing interface of the 3D-RISM solute path:

* ``problem_data(settings)``  -- bgy3d-guile.c:243-385 (defaults, --L is the
  half box edge, h = L/N, rmax = max(L)/2 and nrad = max(N) unless given)
* ``hnc3d_run_solute(solute, solvent, settings, chi=None, restart=None)``
  -- guile/bgy3d.scm:386-388 -> bgy3d-guile.c:1091-1209 run_solute ->
  hnc3d_solute_solve; returns a dict with the same keys as the reference's
  result alist ('free-energy, 'partial-molar-volume-factor, GUV, RESTART)
* ``rism_solvent(solvent, settings)`` -- guile/bgy3d.scm:358 ->
  bgy3d-guile.c:1233-1286; returns {'susceptibility': chi[m][m][nrad]}

The compute path is the gfx950 HIP library only.  There is no CPU fallback:
if the library is missing this module raises, and on a machine without a
GPU the C ABI returns RISM3D_NO_DEVICE, which raises here.
"""
import ctypes
import os

import numpy as np

from .solutes import find_molecule

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(_HERE, "librism3d.so")

CLOSURES = {"HNC": 0, "KH": 1, "PY": 2, **{f"PSE{n}": 3 + n for n in range(8)}}

OK, NONCONVERGED, NAN, BAD_ARG, NO_DEVICE, OOM, UNSUPPORTED, NO_SCRATCH = range(8)
STATUS_NAMES = ["OK", "NONCONVERGED", "NAN", "BAD_ARG", "NO_DEVICE", "OOM", "UNSUPPORTED",
                "NO_SCRATCH"]

SOLVER_MDIIS, SOLVER_PICARD, SOLVER_NEWTON, SOLVER_JAGER, SOLVER_TRIAL = 0, 1, 2, 3, 4
FP64, MIXED, FP32 = 0, 1, 2


class RismError(RuntimeError):
    def __init__(self, status, what="rism3d"):
        self.status = status
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{what}: status {name}")


class Problem(ctypes.Structure):
    """ProblemData, bgy3d.h:228-256."""
    _fields_ = [("N", ctypes.c_int * 3), ("L", ctypes.c_double * 3),
                ("h", ctypes.c_double * 3), ("rmax", ctypes.c_double),
                ("nrad", ctypes.c_int), ("beta", ctypes.c_double),
                ("rho", ctypes.c_double), ("lambda_", ctypes.c_double),
                ("damp", ctypes.c_double), ("max_iter", ctypes.c_int),
                ("norm_tol", ctypes.c_double), ("closure", ctypes.c_int)]


class Site(ctypes.Structure):
    """Site, bgy3d-solutes.h:9-16."""
    _fields_ = [("name", ctypes.c_char * 5), ("x", ctypes.c_double * 3),
                ("sigma", ctypes.c_double), ("epsilon", ctypes.c_double),
                ("charge", ctypes.c_double), ("user", ctypes.c_void_p)]


class Options(ctypes.Structure):
    _fields_ = [("renorm", ctypes.c_int), ("solver", ctypes.c_int),
                ("precision", ctypes.c_int), ("polish_iters", ctypes.c_int),
                ("mixed_tol", ctypes.c_double), ("diis_depth", ctypes.c_int),
                ("diis_step", ctypes.c_double), ("verbose", ctypes.c_int),
                ("device", ctypes.c_int), ("dielectric", ctypes.c_double),
                ("continuation", ctypes.c_int),
                ("chi3d", ctypes.POINTER(ctypes.c_double)),
                ("site_rho", ctypes.POINTER(ctypes.c_double)),
                ("coarse_tol", ctypes.c_double), ("tune", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("mu", ctypes.c_double), ("mu_hnc", ctypes.c_double),
                ("mu_kh", ctypes.c_double), ("mu_gf", ctypes.c_double),
                ("pmv_factor", ctypes.c_double), ("residual", ctypes.c_double),
                ("iterations", ctypes.c_int), ("iterations_fp32", ctypes.c_int),
                ("status", ctypes.c_int),
                ("g", ctypes.POINTER(ctypes.c_double)),
                ("t", ctypes.POINTER(ctypes.c_double)),
                ("gradient", ctypes.POINTER(ctypes.c_double)),
                ("dmu_dsite", ctypes.POINTER(ctypes.c_double)),
                ("adjoint_iterations", ctypes.c_int),
                ("response", ctypes.POINTER(ctypes.c_double)),
                ("info", ctypes.c_void_p),
                ("g_device", ctypes.c_void_p)]


class Info(ctypes.Structure):
    """rism3d_info: the observables print_info prints
    (bgy3d-potential.c:341-443)."""
    _fields_ = [("q_u", ctypes.c_double), ("q_v", ctypes.c_double),
                ("d_n", ctypes.c_double * 3), ("d_u", ctypes.c_double * 3),
                ("d_v", ctypes.c_double * 3),
                ("Q_u", ctypes.c_double * 9), ("Q_v", ctypes.c_double * 9),
                ("U_v_rho_N", ctypes.c_double), ("U_v_rho_u", ctypes.c_double),
                ("rho_v_U_u", ctypes.c_double),
                ("v_sites", ctypes.POINTER(ctypes.c_double)),
                ("potential", ctypes.POINTER(ctypes.c_double))]


EXPORTS = ("rism3d_options_default", "rism3d_solute_solve", "rism3d_batch_solve",
           "rism3d_residual", "rism3d_jacobian", "rism3d_solvent_solve",
           "rism3d_restart_destroy",
           "rism3d_solvent_chi", "rism3d_solvent_chi_mixture", "rism3d_version",
           "rism3d_cache_release",
           "rism3d_profile_start",
           "rism3d_profile_read")

# rism3d_profile_read's kernel ids 0-7 (7 unused)
KERNELS = ("closure_r2c_x", "fft_y_fwd", "zstar", "fft_y_inv", "c2r_x_resid", "closure_r2c_xy",
           "c2r_xy_resid", "unused")

_lib = None


def load():
    """Load the in-tree gfx950 library; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError(f"{LIBRARY} is missing: run `python -m bgy3d_amd.build`")
        lib = ctypes.CDLL(LIBRARY)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        lib.rism3d_options_default.argtypes = [P(Options)]
        lib.rism3d_options_default.restype = None
        lib.rism3d_solute_solve.argtypes = [P(Problem), ctypes.c_int, P(Site), ctypes.c_int,
                                            P(Site), dp, P(Options), P(Result),
                                            P(ctypes.c_void_p)]
        lib.rism3d_solute_solve.restype = ctypes.c_int
        lib.rism3d_batch_solve.argtypes = [P(Problem), ctypes.c_int, P(Site), ctypes.c_int,
                                           P(ctypes.c_int), P(P(Site)), dp, P(Options),
                                           P(Result)]
        lib.rism3d_batch_solve.restype = ctypes.c_int
        lib.rism3d_batch_solve_grids.argtypes = [P(Problem), ctypes.c_int, P(Site), ctypes.c_int,
                                                 P(ctypes.c_int), P(P(Site)), dp, P(Options),
                                                 P(Result)]
        lib.rism3d_batch_solve_grids.restype = ctypes.c_int
        lib.rism3d_residual.argtypes = [P(Problem), ctypes.c_int, P(Site), ctypes.c_int,
                                        P(Site), dp, P(Options), dp, dp]
        lib.rism3d_residual.restype = ctypes.c_int
        lib.rism3d_residual_noalloc.argtypes = lib.rism3d_residual.argtypes
        lib.rism3d_residual_noalloc.restype = ctypes.c_int
        lib.rism3d_jacobian.argtypes = [P(Problem), ctypes.c_int, P(Site), ctypes.c_int,
                                        P(Site), dp, P(Options), dp, dp, ctypes.c_int, dp]
        lib.rism3d_jacobian.restype = ctypes.c_int
        lib.rism3d_solvent_solve.argtypes = [P(Problem), ctypes.c_int, P(Site), P(Options), dp,
                                             dp, dp, dp, P(Result)]
        lib.rism3d_solvent_solve.restype = ctypes.c_int
        lib.rism3d_restart_destroy.argtypes = [ctypes.c_void_p]
        lib.rism3d_restart_destroy.restype = None
        lib.rism3d_solvent_chi.argtypes = [P(Problem), ctypes.c_int, P(Site), ctypes.c_double, dp]
        lib.rism3d_solvent_chi.restype = ctypes.c_int
        lib.rism3d_solvent_chi_mixture.argtypes = [P(Problem), ctypes.c_int, P(Site),
                                                   P(ctypes.c_int), dp, dp]
        lib.rism3d_solvent_chi_mixture.restype = ctypes.c_int
        lib.rism3d_cache_release.argtypes = []
        lib.rism3d_cache_release.restype = ctypes.c_int
        lib.rism3d_version.argtypes = []
        lib.rism3d_version.restype = ctypes.c_char_p
        lib.rism3d_profile_start.argtypes = [ctypes.c_int]
        lib.rism3d_profile_start.restype = ctypes.c_int
        lib.rism3d_profile_read.argtypes = [dp, dp, P(ctypes.c_long)]
        lib.rism3d_profile_read.restype = ctypes.c_int
        _lib = lib
    return _lib


# ------------------------------------------------------------ settings
DEFAULTS = {
    # bgy3d-guile.c:249-270 and guile/bgy3d.scm:323-340
    "beta": 1.6889, "rho": 0.033427745, "lambda": 0.02, "damp-start": 1.0,
    "max-iter": 1500, "norm-tol": 1.0e-7, "L": 10.0, "N": 64, "closure": "HNC",
}


def problem_data(settings):
    """bgy3d-guile.c:243-385: settings dict -> ProblemData."""
    s = dict(DEFAULTS)
    s.update(settings or {})
    pd = Problem()
    n = int(s["N"])
    length = float(s["L"])
    L = [2.0 * length] * 3
    for i, k in enumerate(("sx", "sy", "sz")):
        if k in s:
            L[i] = float(s[k])
    N = [n] * 3
    for i, k in enumerate(("nx", "ny", "nz")):
        if k in s:
            N[i] = int(s[k])
    for i in range(3):
        pd.N[i] = N[i]
        pd.L[i] = L[i]
        pd.h[i] = L[i] / N[i]
    pd.rmax = float(s.get("rmax", max(L) / 2))
    pd.nrad = int(s.get("nrad", max(N)))
    pd.beta = float(s["beta"])
    pd.rho = float(s["rho"])
    pd.lambda_ = float(s["lambda"])
    pd.damp = float(s["damp-start"])
    pd.max_iter = int(s["max-iter"])
    pd.norm_tol = float(s["norm-tol"])
    clo = s["closure"]
    pd.closure = CLOSURES[clo] if isinstance(clo, str) else int(clo)
    return pd


def make_sites(sites):
    """(name, (x,y,z), sigma, epsilon, charge) tuples -> Site array."""
    if isinstance(sites, str):
        sites = find_molecule(sites)
    arr = (Site * len(sites))()
    for i, (name, x, sig, eps, q) in enumerate(sites):
        arr[i].name = name.encode()[:4]
        for d in range(3):
            arr[i].x[d] = float(x[d])
        arr[i].sigma = float(sig)
        arr[i].epsilon = float(eps)
        arr[i].charge = float(q)
    return arr


def make_options(settings=None, **kw):
    s = dict(settings or {})
    s.update(kw)
    opt = Options()
    load().rism3d_options_default(ctypes.byref(opt))
    if s.get("no-renorm"):
        opt.renorm = 0
    if "snes-solver" in s:
        # bgy3d_snes_default's names (bgy3d-snes.c:16-35) plus "mdiis", the
        # default here; an unknown name is an error, as there
        names = {"mdiis": SOLVER_MDIIS, "picard": SOLVER_PICARD, "newton": SOLVER_NEWTON,
                 "jager": SOLVER_JAGER, "trial": SOLVER_TRIAL}
        if s["snes-solver"] not in names:
            raise ValueError(f"No such SNES solver: {s['snes-solver']}")
        opt.solver = names[s["snes-solver"]]
    for key, field in (("precision", "precision"), ("polish-iters", "polish_iters"),
                       ("mixed-tol", "mixed_tol"), ("diis-depth", "diis_depth"),
                       ("diis-step", "diis_step"), ("verbose", "verbose"),
                       ("device", "device"), ("dielectric", "dielectric"),
                       ("continuation", "continuation"), ("coarse-tol", "coarse_tol"),
                       ("tune", "tune")):
        if key in s:
            v = s[key]
            if key == "precision" and isinstance(v, str):
                v = {"fp64": FP64, "mixed": MIXED, "fp32": FP32}[v]
            setattr(opt, field, type(getattr(opt, field))(v))
    return opt


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def hnc3d_run_solute(solute, solvent, settings, chi=None, restart=None, want_t=False,
                     derivatives=False, adjoint=False, response=False, chi3d=None,
                     info=False, site_rho=None):
    """Solve the 3D-RISM solute problem on the GPU.

    Returns a dict like the reference alist: 'free-energy',
    'partial-molar-volume-factor', 'GUV' (g, [m][N2][N1][N0]), plus
    'free-energy-hnc'/'-kh'/'-gf', 'iterations', 'residual', 'RESTART'
    (opaque token, pass back as ``restart=`` and free with
    ``restart_destroy``) if ``restart`` is not None.
    derivatives=True (or the 'derivatives setting, as hnc3d.c:1783 reads it)
    adds 'free-energy-gradient' (n, 3), the reference's TPT gradient;
    response=True (or the 'response setting, hnc3d.c:1784) adds
    'free-energy-response' (n, 3), its linear-response gradient;
    adjoint=True adds 'free-energy-dsite' (n, 6), exact
    d mu / d(x, y, z, q, sigma, epsilon) by the adjoint.
    chi3d: the --solvent-3d susceptibility of hnc3d_run_solvent ('chi-3d',
    complex or its real part), in place of chi (no tau renormalisation).
    info=True adds 'info', the numbers print_info prints
    (bgy3d-potential.c:341-443; see info_dict), and 'medium', the solvent
    potential U_v on the grid (the field of the reference's Context).
    site_rho: per-site densities of a solvent mixture, with chi from
    rism_solvent_mixture."""
    lib = load()
    pd = problem_data(settings)
    opt = make_options(settings)
    sv = make_sites(solvent)
    su = make_sites(solute)
    m = len(sv)
    shape = (m, pd.N[2], pd.N[1], pd.N[0])
    g = np.empty(shape)
    res = Result()
    res.g = _dptr(g)
    t = None
    if want_t:
        t = np.empty(shape)
        res.t = _dptr(t)
    derivatives = derivatives or bool(settings.get("derivatives", False))
    response = response or bool(settings.get("response", False))
    grad = dsite = resp = None
    if response:
        resp = np.full((len(su), 3), np.nan)
        res.response = _dptr(resp)
    if derivatives:
        grad = np.full((len(su), 3), np.nan)
        res.gradient = _dptr(grad)
    if adjoint:
        dsite = np.full((len(su), 6), np.nan)
        res.dmu_dsite = _dptr(dsite)
    rho_arr = None
    if site_rho is not None:
        rho_arr = np.ascontiguousarray(site_rho, dtype=np.float64)
        if rho_arr.shape != (m,):
            raise ValueError(f"site_rho must have {m} entries")
        opt.site_rho = _dptr(rho_arr)
    inf = None
    if info:
        inf = Info()
        vs = np.full(len(su), np.nan)
        pot = np.empty((pd.N[2], pd.N[1], pd.N[0]))
        inf.v_sites = _dptr(vs)
        inf.potential = _dptr(pot)
        res.info = ctypes.cast(ctypes.pointer(inf), ctypes.c_void_p)
    chi_p = None
    if chi3d is not None:
        c3 = np.asarray(chi3d)
        if np.iscomplexobj(c3):
            c3 = c3.real
        elif c3.ndim == 5:
            c3 = c3[..., 0]
        c3 = np.ascontiguousarray(c3, dtype=np.float64)
        want = (m * (m + 1) // 2, pd.N[2], pd.N[1], pd.N[0] // 2 + 1)
        if c3.shape != want:
            raise ValueError(f"chi3d must be {want}, got {c3.shape}")
        opt.chi3d = _dptr(c3)
    if chi is not None:
        chi = np.ascontiguousarray(chi, dtype=np.float64)
        if chi.shape != (m, m, pd.nrad):
            raise ValueError(f"chi must be (m, m, nrad) = {(m, m, pd.nrad)}, got {chi.shape}")
        chi_p = _dptr(chi)
    tok = None
    if restart is not None:
        tok = ctypes.c_void_p(restart if isinstance(restart, int) else restart)
    st = lib.rism3d_solute_solve(ctypes.byref(pd), m, sv, len(su), su, chi_p, ctypes.byref(opt),
                                 ctypes.byref(res), ctypes.byref(tok) if tok is not None else None)
    if st not in (OK, NONCONVERGED):
        raise RismError(st, "rism3d_solute_solve")
    out = {
        "free-energy": res.mu,
        "partial-molar-volume-factor": res.pmv_factor,
        "free-energy-hnc": res.mu_hnc, "free-energy-kh": res.mu_kh,
        "free-energy-gf": res.mu_gf,
        "GUV": g, "iterations": res.iterations, "iterations-fp32": res.iterations_fp32,
        "residual": res.residual, "status": STATUS_NAMES[st],
    }
    if t is not None:
        out["t"] = t
    if grad is not None:
        out["free-energy-gradient"] = grad
    if resp is not None:
        out["free-energy-response"] = resp
    if dsite is not None:
        out["free-energy-dsite"] = dsite
    if dsite is not None or resp is not None:
        out["adjoint-iterations"] = res.adjoint_iterations
    if tok is not None:
        out["RESTART"] = tok.value
    if inf is not None and st == OK:
        out["info"] = info_dict(inf, vs)
        out["medium"] = pot
    return out


def info_dict(inf, vs):
    """rism3d_info as a dict keyed like print_info's lines."""
    return {"q_u": inf.q_u, "q_v": inf.q_v, "d_n": list(inf.d_n), "d_u": list(inf.d_u),
            "d_v": list(inf.d_v), "Q_u": np.array(inf.Q_u).reshape(3, 3),
            "Q_v": np.array(inf.Q_v).reshape(3, 3), "U_v.rho_N": inf.U_v_rho_N,
            "U_v.rho_u": inf.U_v_rho_u, "rho_v.U_u": inf.rho_v_U_u, "v_at_sites": vs}


def hnc3d_run_solvent(solvent, settings, t_init=None, save=None):
    """The 3D solvent-solvent solve on the GPU (hnc3d_solvent_solve,
    hnc3d.c:804-1073).  Returns 'free-energy' (native closure) and
    'free-energy-hnc'/'-kh'/'-gf', 'GVV' (pairs p = i(i+1)/2 + j, j <= i,
    [np][N2][N1][N0]), 'chi-3d' (complex chi - 1 at the corner,
    [np][N2][N1][N0/2+1], the input of hnc3d_run_solute(chi3d=)), 't',
    'iterations', 'residual', 'status'.  save: a directory to write the
    reference's files g%d%d.bin and x%d%d-fft.bin (PETSc binary)."""
    lib = load()
    pd = problem_data(settings)
    opt = make_options(settings)
    sv = make_sites(solvent)
    m = len(sv)
    npair = m * (m + 1) // 2
    shape = (npair, pd.N[2], pd.N[1], pd.N[0])
    g = np.empty(shape)
    t = np.empty(shape)
    chi = np.empty((npair, pd.N[2], pd.N[1], pd.N[0] // 2 + 1, 2))
    ti = None
    if t_init is not None:
        ti = np.ascontiguousarray(t_init, dtype=np.float64)
        if ti.shape != shape:
            raise ValueError(f"t_init must be {shape}, got {ti.shape}")
    res = Result()
    st = lib.rism3d_solvent_solve(ctypes.byref(pd), m, sv, ctypes.byref(opt),
                                  _dptr(ti) if ti is not None else None, _dptr(g), _dptr(chi),
                                  _dptr(t), ctypes.byref(res))
    if st not in (OK, NONCONVERGED):
        raise RismError(st, "rism3d_solvent_solve")
    out = {"free-energy": res.mu, "free-energy-hnc": res.mu_hnc, "free-energy-kh": res.mu_kh,
           "free-energy-gf": res.mu_gf, "GVV": g, "chi-3d": chi[..., 0] + 1j * chi[..., 1],
           "t": t, "iterations": res.iterations, "residual": res.residual,
           "status": STATUS_NAMES[st]}
    if save is not None:
        from . import petsc_io
        pairs = [(i, j) for i in range(m) for j in range(i + 1)]
        for p, (i, j) in enumerate(pairs):
            for a, b in ((i, j), (j, i)):
                petsc_io.write_vec(os.path.join(save, f"g{a}{b}.bin"), g[p])
                petsc_io.write_vec(os.path.join(save, f"x{a}{b}-fft.bin"), chi[p])
    return out


def restart_destroy(token):
    if token:
        load().rism3d_restart_destroy(ctypes.c_void_p(token))


def batch_solve(solutes, solvent, settings, chi, want_g=False, g_device=None):
    """Many solutes, one grid and solvent: rism3d_batch_solve.
    g_device: None, or one device address per solute (a buffer of
    m N2 N1 N0 doubles on settings' device each, e.g. a torch tensor's
    data_ptr()) that receives g on the device."""
    lib = load()
    pd = problem_data(settings)
    opt = make_options(settings)
    sv = make_sites(solvent)
    m = len(sv)
    arrs = [make_sites(s) for s in solutes]
    B = len(arrs)
    ptrs = (ctypes.POINTER(Site) * B)(*[ctypes.cast(a, ctypes.POINTER(Site)) for a in arrs])
    ns = (ctypes.c_int * B)(*[len(a) for a in arrs])
    res = (Result * B)()
    gs = []
    if want_g:
        for b in range(B):
            g = np.empty((m, pd.N[2], pd.N[1], pd.N[0]))
            gs.append(g)
            res[b].g = _dptr(g)
    if g_device is not None:
        if len(g_device) != B:
            raise ValueError("g_device needs one device address per solute")
        for b in range(B):
            res[b].g_device = ctypes.c_void_p(int(g_device[b]))
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    st = lib.rism3d_batch_solve(ctypes.byref(pd), m, sv, B, ns, ptrs, _dptr(chi),
                                ctypes.byref(opt), res)
    if st not in (OK, NONCONVERGED, NAN):
        raise RismError(st, "rism3d_batch_solve")  # whole-call failures only
    out = []
    for b in range(B):
        d = {"free-energy": res[b].mu, "partial-molar-volume-factor": res[b].pmv_factor,
             "iterations": res[b].iterations, "iterations-fp32": res[b].iterations_fp32,
             "residual": res[b].residual, "status": STATUS_NAMES[res[b].status]}
        if want_g:
            d["GUV"] = gs[b]
        out.append(d)
    return out


def batch_solve_grids(solutes, solvent, settings, chi, want_g=False):
    """Solutes on uneven grids in one call: rism3d_batch_solve_grids.
    settings: one settings dict per solute (its own N, L, ...; options such
    as precision are taken from the first).  Solutes with equal problems
    form a size class; the classes run concurrently on their own streams."""
    lib = load()
    if len(settings) != len(solutes):
        raise ValueError("one settings dict per solute")
    B = len(solutes)
    pds = (Problem * B)()
    for b, st in enumerate(settings):
        pds[b] = problem_data(st)
    opt = make_options(settings[0])
    sv = make_sites(solvent)
    m = len(sv)
    arrs = [make_sites(s) for s in solutes]
    ptrs = (ctypes.POINTER(Site) * B)(*[ctypes.cast(a, ctypes.POINTER(Site)) for a in arrs])
    ns = (ctypes.c_int * B)(*[len(a) for a in arrs])
    res = (Result * B)()
    gs = []
    if want_g:
        for b in range(B):
            N = pds[b].N
            g = np.empty((m, N[2], N[1], N[0]))
            gs.append(g)
            res[b].g = _dptr(g)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    st = lib.rism3d_batch_solve_grids(pds, m, sv, B, ns, ptrs, _dptr(chi), ctypes.byref(opt), res)
    if st not in (OK, NONCONVERGED, NAN):
        raise RismError(st, "rism3d_batch_solve_grids")
    out = []
    for b in range(B):
        d = {"free-energy": res[b].mu, "partial-molar-volume-factor": res[b].pmv_factor,
             "iterations": res[b].iterations, "iterations-fp32": res[b].iterations_fp32,
             "residual": res[b].residual, "status": STATUS_NAMES[res[b].status]}
        if want_g:
            d["GUV"] = gs[b]
        out.append(d)
    return out


def residual(solute, solvent, settings, chi, t, noalloc=False):
    """One evaluation of F(t) = (chi - 1) * c(t) - t (iterate_t1).
    noalloc=True calls rism3d_residual_noalloc (no device allocation, fixed
    scratch of RISM3D_NOALLOC_SCRATCH_MB per device)."""
    lib = load()
    pd = problem_data(settings)
    opt = make_options(settings)
    sv = make_sites(solvent)
    su = make_sites(solute)
    t = np.ascontiguousarray(t, dtype=np.float64)
    r = np.empty_like(t)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    fn = lib.rism3d_residual_noalloc if noalloc else lib.rism3d_residual
    st = fn(ctypes.byref(pd), len(sv), sv, len(su), su, _dptr(chi), ctypes.byref(opt), _dptr(t),
            _dptr(r))
    if st != OK:
        raise RismError(st, fn.__name__)
    return r


def jacobian(solute, solvent, settings, chi, t, dt, flag=False):
    """One application of the Jacobian at t (jacobian_t1, hnc3d.c:1290-1362):
    (chi - 1) * (dc/dt dt) - dt, or with flag its flag form (chi - 1) *
    (dc/dt dt + dt)."""
    lib = load()
    pd = problem_data(settings)
    opt = make_options(settings)
    sv = make_sites(solvent)
    su = make_sites(solute)
    t = np.ascontiguousarray(t, dtype=np.float64)
    dt = np.ascontiguousarray(dt, dtype=np.float64)
    out = np.empty_like(t)
    chi = np.ascontiguousarray(chi, dtype=np.float64)
    st = lib.rism3d_jacobian(ctypes.byref(pd), len(sv), sv, len(su), su, _dptr(chi),
                             ctypes.byref(opt), _dptr(t), _dptr(dt), int(bool(flag)), _dptr(out))
    if st != OK:
        raise RismError(st, "rism3d_jacobian")
    return out


def rism_solvent(solvent, settings):
    """1D DRISM susceptibility (rism_solvent, rism.f90:57-102)."""
    lib = load()
    pd = problem_data(settings)
    sv = make_sites(solvent)
    m = len(sv)
    chi = np.empty((m, m, pd.nrad))
    eps = float((settings or {}).get("dielectric", 0.0))
    st = lib.rism3d_solvent_chi(ctypes.byref(pd), m, sv, eps, _dptr(chi))
    if st != OK:
        raise RismError(st, "rism3d_solvent_chi")
    return {"susceptibility": chi}


def rism_solvent_mixture(solvent, molecule, rho, settings):
    """1D XRISM susceptibility of a solvent mixture (no counterpart in the
    reference): `molecule[i]` the molecule of site i, `rho[i]` its number
    density.  Returns {'susceptibility': chi} with
    chi_ij = w_ij + rho_i h_ij, for hnc3d_run_solute(chi=, site_rho=rho)."""
    lib = load()
    pd = problem_data(settings)
    sv = make_sites(solvent)
    m = len(sv)
    mol = np.ascontiguousarray(molecule, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    if mol.shape != (m,) or rho.shape != (m,):
        raise ValueError(f"molecule and rho must have {m} entries")
    chi = np.empty((m, m, pd.nrad))
    st = lib.rism3d_solvent_chi_mixture(ctypes.byref(pd), m, sv,
                                        mol.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        _dptr(rho), _dptr(chi))
    if st != OK:
        raise RismError(st, "rism3d_solvent_chi_mixture")
    return {"susceptibility": chi}


def cache_release():
    """Free the device buffers the library keeps between solves."""
    load().rism3d_cache_release()


def profile_start(device=0):
    st = load().rism3d_profile_start(device)
    if st != OK:
        raise RismError(st, "rism3d_profile_start")


def profile_read():
    """{kernel<type>: (ms, algorithmic bytes, launches)} since profile_start,
    one entry per residual kernel and working precision, e.g.
    "closure_r2c_x<float>"; instantiations never launched are left out."""
    nk = len(KERNELS)
    ms = np.zeros(2 * nk)
    by = np.zeros(2 * nk)
    n = (ctypes.c_long * (2 * nk))()
    load().rism3d_profile_read(_dptr(ms), _dptr(by), n)
    out = {}
    for slot, typ in ((0, "double"), (nk, "float")):
        for i, k in enumerate(KERNELS):
            if n[slot + i]:
                out[f"{k}<{typ}>"] = (float(ms[slot + i]), float(by[slot + i]), int(n[slot + i]))
    return out
