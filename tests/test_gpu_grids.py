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
