"""Pin oracle/closures.py (closures.f90:107-653) by closed forms.

Each check is chosen so that a plausible slip (wrong branch boundary,
dropped -t, wrong factorial, sign) fails it: scalar math-module evaluation
at every branch, KH == PSE1 (x != 0), PSE-n -> HNC as n grows, and central
finite differences of closure() against closure1().
"""
import math

import numpy as np
import pytest

from oracle import closures as C

BETA = 1.6889


def scalar_closure(method, beta, v, t):
    bv = beta * v
    x = -bv + t
    if method == C.HNC:
        return math.expm1(x) - t
    if method == C.KH:
        return math.expm1(x) - t if x <= 0 else -bv
    if method == C.PY:
        return math.exp(-bv) * (1 + t) - 1 - t
    n = method - C.PSE0
    if x < 0:
        return math.expm1(x) - t
    return sum(x ** k / math.factorial(k) for k in range(1, n + 1)) - t


GRID = [(v, t) for v in (-3.0, -0.5, 0.0, 0.3, 2.0, 50.0) for t in (-2.0, -0.1, 0.0, 0.4, 3.0)]


@pytest.mark.parametrize("method", list(range(11)))
def test_closure_scalar(method):
    v = np.array([g[0] for g in GRID])
    t = np.array([g[1] for g in GRID])
    got = C.closure(method, BETA, v, t)
    want = [scalar_closure(method, BETA, a, b) for a, b in GRID]
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)


def test_kh_equals_pse1():
    rng = np.random.default_rng(1)
    v = rng.uniform(-2, 2, 1000)
    t = rng.uniform(-2, 2, 1000)
    np.testing.assert_allclose(C.closure(C.KH, BETA, v, t), C.closure(C.PSE0 + 1, BETA, v, t),
                               rtol=1e-14, atol=1e-15)


def test_pse_to_hnc():
    x = np.linspace(0.0, 0.5, 11)
    t = np.zeros_like(x)
    v = -x / BETA
    hnc = C.closure(C.HNC, BETA, v, t)
    pse7 = C.closure(C.PSE7, BETA, v, t)
    assert np.max(np.abs(hnc - pse7)) < 0.5 ** 8 / math.factorial(8) * 1.1


@pytest.mark.parametrize("method", [C.HNC, C.KH, C.PSE0 + 2, C.PSE0 + 3, C.PSE7])
def test_closure1_finite_difference(method):
    rng = np.random.default_rng(2)
    v = rng.uniform(-1, 1, 400)
    t = rng.uniform(-1, 1, 400)
    keep = np.abs(-BETA * v + t) > 1e-3    # away from the branch point
    v, t = v[keep], t[keep]
    h = 1e-6
    dt = np.ones_like(t)
    fd = (C.closure(method, BETA, v, t + h) - C.closure(method, BETA, v, t - h)) / (2 * h)
    # closure1 is d(c + t)/dt - dt ... c = f(x) - t, so dc/dt = f'(x) - 1
    np.testing.assert_allclose(C.closure1(method, BETA, v, t, dt), fd, rtol=1e-6, atol=1e-8)


def test_forms():
    rng = np.random.default_rng(3)
    x, h, c, cl = (rng.uniform(-1, 1, 500) for _ in range(4))
    base = -c - h * (c + cl) / 2
    np.testing.assert_allclose(C.form(C.HNC, x, h, c, cl), base + h * h / 2)
    np.testing.assert_allclose(C.form(C.PY, x, h, c, cl), base)
    kh = C.form(C.KH, x, h, c, cl)
    np.testing.assert_allclose(kh[h < 0], (base + h * h / 2)[h < 0])
    np.testing.assert_allclose(kh[h >= 0], base[h >= 0])
    # PSE1 with h = x (the closure relation where x >= 0) equals KH
    hp = np.abs(h)
    np.testing.assert_allclose(C.form(C.PSE0 + 1, hp, hp, c, cl), C.form(C.KH, hp, hp, c, cl),
                               atol=1e-15)
    # PSE3 tail term x^4/4!
    p3 = C.form(C.PSE0 + 3, x, hp, c, cl)
    np.testing.assert_allclose(p3, -c - hp * (c + cl) / 2 + hp ** 2 / 2 - x ** 4 / 24)


def test_form1_finite_difference():
    rng = np.random.default_rng(4)
    x, h, c, cl = (rng.uniform(-1, 1, 300) for _ in range(4))
    dx, dh, dc = (rng.uniform(-1, 1, 300) for _ in range(3))
    e = 1e-6
    for method in (C.HNC, C.KH, C.PY, C.PSE0 + 3):
        keep = np.abs(h) > 1e-3
        fd = (C.form(method, x + e * dx, h + e * dh, c + e * dc, cl)
              - C.form(method, x - e * dx, h - e * dh, c - e * dc, cl)) / (2 * e)
        got = C.form1(method, x, dx, h, dh, c, dc, cl)
        np.testing.assert_allclose(got[keep], fd[keep], rtol=1e-6, atol=1e-8)
