"""The C ABI without a GPU: the library loads, exports every entry point
include/rism3d.h declares, its structs have the layout the header defines
(checked against gcc on the header itself), argument validation runs before
any device call, and compute calls fail loudly (NO_DEVICE) without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rism3d.h")


def _R():
    from bgy3d_amd import rism3d as R
    return R


def test_exports_every_declared_symbol():
    R = _R()
    lib = R.load()
    text = open(HEADER).read()
    declared = sorted(set(re.findall(r"\b(rism3d_[a-z_0-9]+)\s*\(", text)))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in rism3d.h but not exported"
    assert set(R.EXPORTS) <= set(declared)


def test_struct_layout_matches_header(tmp_path):
    R = _R()
    prog = tmp_path / "layout.c"
    fields = {
        "rism3d_problem": [f for f, _ in R.Problem._fields_],
        "rism3d_site": [f for f, _ in R.Site._fields_],
        "rism3d_options": [f for f, _ in R.Options._fields_],
        "rism3d_result": [f for f, _ in R.Result._fields_],
        "rism3d_info": [f for f, _ in R.Info._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            cf = "lambda" if f == "lambda_" else f
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {cf}));')
    lines.append("return 0;}")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(prog), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.split("\n") if l)
    py = {"rism3d_problem": R.Problem, "rism3d_site": R.Site, "rism3d_options": R.Options,
          "rism3d_result": R.Result, "rism3d_info": R.Info}
    for st, cls in py.items():
        assert int(out[st]) == ctypes.sizeof(cls), st
        for f, _ in cls._fields_:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_problem_layout_is_reference_problem_data():
    """ProblemData (bgy3d.h:228-256): int N[3]; double L[3], h[3], rmax;
    int nrad; double beta, rho, lambda, damp; int max_iter; double norm_tol;
    ClosureEnum closure -- same order and types."""
    R = _R()
    names = [f for f, _ in R.Problem._fields_]
    assert names == ["N", "L", "h", "rmax", "nrad", "beta", "rho", "lambda_", "damp",
                     "max_iter", "norm_tol", "closure"]


def test_problem_data_semantics():
    R = _R()
    pd = R.problem_data({"L": 10, "N": 96})
    assert list(pd.N) == [96] * 3 and list(pd.L) == [20.0] * 3      # --L is the half edge
    assert abs(pd.h[0] - 20.0 / 96) < 1e-15
    assert pd.rmax == 10.0 and pd.nrad == 96                         # bgy3d-guile.c:328-331
    assert pd.closure == 0 and pd.max_iter == 1500
    pd = R.problem_data({"L": 10, "N": 64, "closure": "PSE3", "nrad": 4096, "rmax": 160.0})
    assert pd.closure == 6 and pd.nrad == 4096 and pd.rmax == 160.0


def _chi(m=3, nrad=1536):
    return np.tile(np.eye(m)[:, :, None], (1, 1, nrad))


def _gpu_present():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def test_validation_before_device():
    R = _R()
    W = "water, PR-SPC/E"
    base = {"L": 10, "N": 32, "rmax": 40.0, "nrad": 1536}
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solute(W, W, dict(base, N=100), chi=_chi())
    assert e.value.status == R.UNSUPPORTED
    with pytest.raises(R.RismError) as e:  # k table too short (interp asserts)
        R.hnc3d_run_solute(W, W, dict(base, nrad=16), chi=_chi(nrad=16))
    assert e.value.status == R.BAD_ARG
    with pytest.raises(ValueError):
        R.hnc3d_run_solute(W, W, base, chi=_chi(nrad=100))
    # the reference's Newton / trial / Jager solvers run in FP64 only, and
    # Newton needs the closure differential PY lacks (closures.f90:131-153)
    for solver in ("newton", "jager", "trial"):
        with pytest.raises(R.RismError) as e:
            R.hnc3d_run_solute(W, W, dict(base, **{"snes-solver": solver, "precision": "mixed"}),
                               chi=_chi())
        assert e.value.status == R.UNSUPPORTED
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solute(W, W, dict(base, closure="PY", **{"snes-solver": "newton"}),
                           chi=_chi())
    assert e.value.status == R.UNSUPPORTED
    with pytest.raises(R.RismError) as e:
        R.jacobian(W, W, dict(base, closure="PY"), _chi(), np.zeros((3, 32, 32, 32)),
                   np.zeros((3, 32, 32, 32)))
    assert e.value.status == R.UNSUPPORTED


def test_solvent_solve_validation():
    """rism3d_solvent_solve: m <= 8 sites (the per-k OZ is solved by
    elimination per k point), the site-site pairs of a z line tile within
    64 KB of LDS, cubic boxes only (omega_ij(k) takes L[0] for every axis,
    as bgy3d-pure.c:236-267), FP64 MDIIS or Picard."""
    R = _R()
    base = {"L": 10, "N": 32}
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solvent("hexane", base)            # 20 sites
    assert e.value.status == R.BAD_ARG
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solvent("methanol", dict(base, N=128))  # 21 pairs x 129 points x 2 lines
    assert e.value.status == R.UNSUPPORTED
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solvent("water", dict(base, precision="mixed"))
    assert e.value.status == R.UNSUPPORTED
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solvent("water", dict(base, **{"snes-solver": "newton"}))
    assert e.value.status == R.UNSUPPORTED
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solvent("water", dict(base, N=100))
    assert e.value.status == R.UNSUPPORTED
    with pytest.raises(ValueError):
        R.hnc3d_run_solvent("water", base, t_init=np.zeros((3, 32, 32, 32)))


@pytest.mark.skipif(_gpu_present(), reason="checks the no-GPU error path")
def test_no_device_fails_loudly():
    R = _R()
    W = "water, PR-SPC/E"
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solute(W, W, {"L": 10, "N": 32, "rmax": 40.0, "nrad": 1536}, chi=_chi())
    assert e.value.status == R.NO_DEVICE


def test_petsc_vec_format(tmp_path):
    """bgy3d_vec_save layout: int32 class id 1211214, int32 length, then
    big-endian doubles in natural order."""
    from bgy3d_amd import petsc_io
    a = np.arange(24, dtype=np.float64).reshape(2, 3, 4) * 0.5
    path = tmp_path / "g0.bin"
    petsc_io.write_vec(path, a)
    raw = open(path, "rb").read()
    assert raw[:4] == (1211214).to_bytes(4, "big") and raw[4:8] == (24).to_bytes(4, "big")
    assert np.frombuffer(raw[8:], dtype=">f8")[5] == 2.5
    assert np.array_equal(petsc_io.read_vec(path, a.shape), a)


def test_mixture_rejects_gradients():
    """options.site_rho: the gradients are built for one density, and the
    mixture's chi must come from the caller."""
    R = _R()
    W = "water, PR-SPC/E"
    base = {"L": 10, "N": 32, "rmax": 40.0, "nrad": 1536}
    rho = [0.0333295] * 3
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solute(W, W, base, chi=_chi(), site_rho=rho, adjoint=True)
    assert e.value.status == R.UNSUPPORTED
    with pytest.raises(R.RismError) as e:
        R.hnc3d_run_solute(W, W, base, site_rho=rho)
    assert e.value.status == R.BAD_ARG


def test_cache_release_without_work():
    """rism3d_cache_release with an empty cache (no device needed)."""
    R = _R()
    assert R.load().rism3d_cache_release() == R.OK


def test_unknown_snes_solver_is_an_error():
    """bgy3d_snes_default (bgy3d-snes.c:16-35) exits on an unknown solver
    name; the options mapping refuses it before any device work."""
    from bgy3d_amd import rism3d as R
    for name in ("newton", "picard", "jager", "trial", "mdiis"):
        R.make_options({"snes-solver": name})
    with pytest.raises(ValueError, match="No such SNES solver"):
        R.make_options({"snes-solver": "broyden"})


def test_grids_reject_mixed_radial_grids():
    """rism3d_batch_solve_grids shares one chi: problems on different radial
    grids are refused before any device work."""
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    chi = np.zeros((3, 3, 1536))
    a = {"N": 32, "L": 8.0, "rmax": 40.0, "nrad": 1536}
    b = dict(a, rmax=39.0)
    with pytest.raises(R.RismError) as e:
        R.batch_solve_grids([S.solute(0), S.solute(1)], "water, PR-SPC/E", [a, b], chi)
    assert e.value.status == R.BAD_ARG
