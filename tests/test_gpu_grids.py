"""rism3d_batch_solve_grids: solutes on uneven grids in one call, the size
classes solved concurrently on their own streams, against each class
solved alone by rism3d_batch_solve (same kernels, same inputs,
deterministic reductions: bit-identical mu and iteration counts), in FP64
and mixed precision."""
import pytest

from conftest import README

pytestmark = pytest.mark.gpu
W = "water, PR-SPC/E"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_grids_equal_batches(chi_w, precision):
    print(f"case batch_solve_grids {precision}", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    sols = [S.solute(s) for s in range(6)]
    grids = [(32, 8.0), (48, 9.0), (32, 8.0), (64, 10.0), (48, 9.0), (32, 8.0)]

    def st(N, L):
        s = dict(README, N=N, L=L, rmax=40.0, nrad=1536, closure="KH", precision=precision,
                 **{"norm-tol": 1e-9, "max-iter": 3000})
        return s

    out = R.batch_solve_grids(sols, W, [st(N, L) for N, L in grids], chi_w)
    for N, L in sorted(set(grids)):
        idx = [i for i, g in enumerate(grids) if g == (N, L)]
        ref = R.batch_solve([sols[i] for i in idx], W, st(N, L), chi_w)
        for i, r in zip(idx, ref):
            print(f"  solute {i} {N}^3: mu {out[i]['free-energy']!r} alone {r['free-energy']!r} "
                  f"iters {out[i]['iterations']} {r['iterations']} {out[i]['status']}", flush=True)
            assert out[i]["status"] == "OK" and r["status"] == "OK"
            assert out[i]["free-energy"] == r["free-energy"]
            assert out[i]["iterations"] == r["iterations"]


def test_tuner_records_choices_and_keeps_results(chi_w, tmp_path, monkeypatch):
    """options.tune: each working precision's launch configuration timed on
    the batch and the fastest kept.  FP64 candidates differ only in the
    z-pass block order (the same arithmetic per point: bit-identical mu);
    mixed ones also in the FP32 chain layout (another rounding order, the
    same fixed point).  The choices are appended to RISM3D_TUNE_FILE."""
    print("case tuner", flush=True)
    from bgy3d_amd import rism3d as R
    from bgy3d_amd import synthetic as S
    f = tmp_path / "tune.txt"
    monkeypatch.setenv("RISM3D_TUNE_FILE", str(f))
    sols = [S.solute(s) for s in range(4)]
    base = dict(README, N=64, L=10.0, rmax=40.0, nrad=1536, closure="KH",
                **{"norm-tol": 1e-9, "max-iter": 3000})
    a = R.batch_solve(sols, W, base, chi_w)
    b = R.batch_solve(sols, W, dict(base, tune=1), chi_w)
    for x, y in zip(a, b):
        assert x["status"] == "OK" and y["status"] == "OK"
        assert x["free-energy"] == y["free-energy"]
    am = R.batch_solve(sols, W, dict(base, precision="mixed"), chi_w)
    bm = R.batch_solve(sols, W, dict(base, precision="mixed", tune=1), chi_w)
    for x, y, z in zip(am, bm, a):
        print(f"  mixed {x['free-energy']:.8f} tuned {y['free-energy']:.8f} fp64 {z['free-energy']:.8f}",
              flush=True)
        assert x["status"] == "OK" and y["status"] == "OK"
        assert abs(y["free-energy"] - z["free-energy"]) < 0.01
    lines = f.read_text().split("\n")
    print("  " + "\n  ".join(lines), flush=True)
    keys = [ln.split()[0] for ln in lines if ln.strip()]
    assert "64x64x64,m3,B4,fp64" in keys and "64x64x64,m3,B4,fp32" in keys
