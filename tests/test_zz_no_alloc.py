"""rism3d_residual_noalloc (no device allocation, fixed scratch) against
rism3d_residual through the C ABI: two calls in a row, the second on a
smaller grid, in FP64 and in the FP32 (mixed) kernels.  The two entry points
run the same kernels on the same inputs, so the residuals must agree bit
for bit.  Also: device memory stays where it was across calls (the
library's own buffers would be tens of MB), and a problem larger than the
scratch is refused with RISM3D_NO_SCRATCH."""
import numpy as np
import pytest

from conftest import CHI_W, README, load_chi

pytestmark = pytest.mark.gpu

W = "water, PR-SPC/E"
SOLUTE = [("C", (0.4, -0.3, 0.2), 3.2, 0.15, 0.6), ("O", (-0.9, 0.6, -0.4), 3.0, 0.17, -0.8),
          ("H", (1.6, 0.3, 0.9), 1.0, 0.05, 0.2)]


def _settings(N, precision="fp64", **kw):
    s = dict(README, N=N, L=10.0, rmax=40.0, nrad=1536, closure="KH", device=0,
             precision=precision)
    s.update(kw)
    return s


def _t(shape, seed):
    rng = np.random.default_rng(seed)
    m, n2, n1, n0 = shape
    z, y, x = np.meshgrid(*(np.linspace(-1, 1, n, endpoint=False) for n in (n2, n1, n0)),
                          indexing="ij")
    r2 = x * x + y * y + z * z
    t = np.stack([rng.uniform(-1, 1) * np.exp(-r2 / rng.uniform(0.05, 0.3)) for _ in range(m)])
    return t + 0.01 * rng.standard_normal(t.shape)


def _free_bytes():
    """hipMemGetInfo of the HIP runtime the library itself loaded."""
    import ctypes
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    # torch, when imported first, brings the runtime the library then binds
    # to; when both are loaded, the library's is the one outside torch/lib
    ours = [p for p in paths if "/torch/" not in p] if len(paths) > 1 else list(paths)
    assert len(ours) == 1, paths
    path = ours[0]
    hip = ctypes.CDLL(path)
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipSetDevice(0) == 0
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_residual(precision):
    from bgy3d_amd import rism3d as R
    chi = load_chi(CHI_W)
    cases = [(_settings(64, precision), 1), (_settings(32, precision, nz=48), 2)]
    # first call on the device: scratch address and stream set up
    s0, seed0 = cases[0]
    t0 = _t((3, 64, 64, 64), seed0)
    R.residual(SOLUTE, W, s0, chi, t0, noalloc=True)
    before = _free_bytes()
    for s, seed in cases:
        shape = (3, s.get("nz", s["N"]), s["N"], s["N"])
        print(f"noalloc {precision} grid {shape[1:]} seed {seed}", flush=True)
        t = _t(shape, seed)
        r_na = R.residual(SOLUTE, W, s, chi, t, noalloc=True)
        after = _free_bytes()
        r_ref = R.residual(SOLUTE, W, s, chi, t)
        assert np.all(np.isfinite(r_na))
        np.testing.assert_array_equal(r_na, r_ref)
        # no device buffers were taken by the allocation-free calls
        assert abs(after - before) < (4 << 20), (before, after)
        before = _free_bytes()


def test_noalloc_refuses_too_large():
    from bgy3d_amd import rism3d as R
    chi = load_chi(CHI_W)
    t = np.zeros((3, 96, 96, 96))
    with pytest.raises(R.RismError) as e:
        R.residual(SOLUTE, W, _settings(96), chi, t, noalloc=True)
    assert e.value.status == R.NO_SCRATCH
