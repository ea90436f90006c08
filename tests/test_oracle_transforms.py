"""Pin the oracle's transforms and grid conventions by mathematics.

* DST-IV (fft.f90 rism_dst = FFTW RODFT11) against the O(n^2) defining sum.
* fourier_rows (fft.f90:121-170) against the closed-form 3D FT of a
  Gaussian, and the forward/backward pair against identity.
* r2c/c2r: plane-wave known answers and the round trip x N^3
  (bgy3d-fft.c:16-72 idea), incl. FFTW's dropping of Im at DC/Nyquist.
* s(k) = (-1)^(kx+ky+kz): shifting a function by L/2 along every axis.
* interp (bgy3d-vec.c:23-46): exact on linear tables, index clamping rules.
"""
import math

import numpy as np

from oracle import rism1d, rism3d


def test_dst4_definition():
    rng = np.random.default_rng(0)
    for n in (7, 16, 48):
        x = rng.standard_normal(n)
        j = np.arange(n)
        k = j[:, None]
        want = 2 * np.sum(x[None, :] * np.sin(math.pi * (2 * j[None, :] + 1) * (2 * k + 1) / (4 * n)), axis=1)
        np.testing.assert_allclose(rism1d.dst4(x), want, rtol=1e-12, atol=1e-12)


def test_radial_ft_gaussian():
    rmax, nrad = 40.0, 2048
    r, dr, k, dk = rism1d.mkgrid(rmax, nrad)
    a = 0.7
    f = np.exp(-a * r * r)
    g = rism1d.fourier_rows(f[None, None, :])[0, 0] * dr ** 3
    want = (math.pi / a) ** 1.5 * np.exp(-k * k / (4 * a))
    np.testing.assert_allclose(g, want, atol=1e-12)
    back = rism1d.fourier_rows(g[None, None, :])[0, 0] * dk ** 3 / (2 * math.pi) ** 3
    np.testing.assert_allclose(back, f, atol=1e-12)
    # integrate(): 4 pi int f r^2 dr == FT at k = 0
    assert abs(rism1d.integrate(f) * dr ** 3 - (math.pi / a) ** 1.5) < 1e-10


def test_plane_wave_r2c_c2r():
    pd = rism3d.Problem((8, 6, 10), (2.0, 3.0, 4.0), 1.0, 1.0)
    N0, N1, N2 = pd.N
    z, y, x = np.meshgrid(np.arange(N2), np.arange(N1), np.arange(N0), indexing="ij")
    a = np.cos(2 * math.pi * (2 * x / N0 + 1 * y / N1 + 3 * z / N2))
    F = rism3d.rfft(pd, a)
    P = N0 * N1 * N2
    assert abs(F[3, 1, 2] - P / 2) < 1e-9 and abs(F[N2 - 3, N1 - 1, 2]) < 1e-9
    F[np.abs(F) < 1e-9] = 0
    np.testing.assert_allclose(rism3d.irfft(pd, F), P * a, atol=1e-9)
    # c2r drops Im at DC and Nyquist (FFTW/pocketfft semantics)
    G = np.zeros_like(F)
    G[0, 0, 0] = 1 + 5j
    G[0, 0, N0 // 2] = 2 - 3j
    out = rism3d.irfft(pd, G)
    want = 1 + 2 * np.cos(math.pi * x)
    np.testing.assert_allclose(out, want, atol=1e-12)


def test_phase_sign_is_half_box_shift():
    pd = rism3d.Problem(16, 10.0, 1.0, 1.0)
    x, y, z = rism3d.r_axes(pd)
    centre = np.exp(-(z[:, None, None] ** 2 + y[None, :, None] ** 2 + x[None, None, :] ** 2))
    # the same function moved by L/2 on every axis (periodic image at the corner)
    corner = np.roll(centre, (8, 8, 8), axis=(0, 1, 2))
    np.testing.assert_allclose(rism3d.rfft(pd, corner) * rism3d.phase_sign(pd),
                               rism3d.rfft(pd, centre), atol=1e-10)


def test_interp_rules():
    dx = 0.5
    tab = 3.0 + 2.0 * (np.arange(20) + 0.5) * dx           # f(x) = 3 + 2x
    xs = np.array([0.0, 0.1, 1.3, 5.0, 9.7])
    np.testing.assert_allclose(rism3d.interp(xs, tab, dx), 3 + 2 * xs, atol=1e-12)
    # one-sided differences at the ends: i <= 1 uses tab[i+1]-tab[i]
    q = (np.arange(20) + 0.5) ** 2
    got = rism3d.interp(np.array([0.0]), q, 1.0)[0]
    assert got == q[0] + (q[1] - q[0]) * (-0.5)
