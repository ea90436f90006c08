"""The coarse-start checker (tests/trig_upsample.py) against what the
mathematics fixes: a field band-limited below the coarse Nyquist frequency,
sampled on N, is reproduced exactly on 2N; every coarse node is kept; a
pure Nyquist cosine goes to the cosine between the nodes.  No GPU."""
import numpy as np

from trig_upsample import upsample


def _grid(n, L):
    return np.arange(n) * (L / n) - L / 2


def test_band_limited_field_is_exact():
    rng = np.random.default_rng(0)
    L = (7.0, 9.0, 11.0)
    nc = (8, 12, 16)  # N0, N1, N2 of the coarse grid
    def field(n0, n1, n2):
        x, y, z = _grid(n0, L[0]), _grid(n1, L[1]), _grid(n2, L[2])
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        f = np.zeros(Z.shape)
        for _ in range(6):
            kx, ky, kz = rng.integers(-3, 4), rng.integers(-5, 6), rng.integers(-7, 8)
            a, ph = rng.normal(), rng.uniform(0, 2 * np.pi)
            f += a * np.cos(2 * np.pi * (kx * X / L[0] + ky * Y / L[1] + kz * Z / L[2]) + ph)
        return f
    state = rng.bit_generator.state
    coarse = field(*nc)
    rng.bit_generator.state = state
    fine = field(*(2 * n for n in nc))
    up = upsample(coarse[None])[0]
    assert up.shape == fine.shape
    assert np.max(np.abs(up - fine)) < 1e-12 * np.max(np.abs(fine)) * 100


def test_coarse_nodes_kept_and_nyquist_split():
    rng = np.random.default_rng(1)
    t = rng.normal(size=(2, 6, 8, 10))
    up = upsample(t)
    assert np.max(np.abs(up[:, ::2, ::2, ::2] - t)) < 1e-12
    # cos(pi i) on the coarse x axis -> cos(pi i / 2) on the fine one
    n = 8
    c = np.cos(np.pi * np.arange(n))[None, None, None, :] * np.ones((1, 4, 4, n))
    f = upsample(c)
    want = np.cos(np.pi * np.arange(2 * n) / 2)[None, None, None, :] * np.ones((1, 8, 8, 2 * n))
    assert np.max(np.abs(f - want)) < 1e-12
