"""Closures, closure differentials and chemical-potential densities (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates reference ``closures.f90``:

* closure      closures.f90:107-128  (closure_hnc 248, closure_kh 279,
                                      closure_py 314, closure_pse 358, pse 392,
                                      exps1 191)
* closure1     closures.f90:131-153  (closure_hnc1 259, closure_kh1 299,
                                      closure_pse1 375)
* form         closures.f90:574-612  (chempot density, Kast-Kloss forms)
* form1        closures.f90:615-653

Closure numbering follows bgy3d.h:211-223 / foreign.f90:42-57:
HNC=0, KH=1, PY=2, PSE0..PSE7 = 3..10.

Pinning: closed forms checked in tests/test_oracle_closures.py against
(i) a direct ``math.expm1``/``math.factorial`` scalar evaluation at every
branch, (ii) the identity KH == PSE1 away from x == 0, (iii) PSE-n -> HNC
as n grows for x > 0 small, (iv) central finite differences of closure()
against closure1().
"""
import math

import numpy as np

HNC, KH, PY = 0, 1, 2
PSE0 = 3
PSE7 = 10
NAMES = {"HNC": 0, "KH": 1, "PY": 2, **{f"PSE{n}": 3 + n for n in range(8)}}


def order(method):
    """closures.f90:156-174: HNC -> infinite, KH -> 1, PSEn -> n."""
    if method == HNC:
        return 10**9
    if method == KH:
        return 1
    if PSE0 <= method <= PSE7:
        return method - PSE0
    return -1


def exps1(n, x):
    """Truncated exponential series sum_{k=1..n} x^k/k! by Horner
    (closures.f90:191-236)."""
    x = np.asarray(x, dtype=np.float64)
    c = 1.0
    y = np.zeros_like(x)
    for k in range(n, 0, -1):
        y = x * y + c
        c = c * k
    return (x / c) * y


def pse(n, x):
    """closures.f90:392-418: expm1(x) for x < 0, exps1(n, x) otherwise."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(over="ignore"):
        e = np.expm1(np.minimum(x, 0.0))
    return np.where(x < 0.0, e, exps1(n, np.maximum(x, 0.0)))


def closure(method, beta, v, t):
    """c = closure(beta*v, t), elementwise (closures.f90:107-128)."""
    bv = beta * np.asarray(v, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    x = -bv + t
    if method == HNC:
        with np.errstate(over="ignore"):
            return np.expm1(x) - t
    if method == KH:
        return np.where(x <= 0.0, np.expm1(np.minimum(x, 0.0)) - t, -bv)
    if method == PY:
        with np.errstate(over="ignore"):
            return np.exp(-bv) * (1.0 + t) - 1.0 - t
    if PSE0 <= method <= PSE7:
        return pse(order(method), x) - t
    raise ValueError(f"no such closure {method}")


def closure1(method, beta, v, t, dt):
    """dc = closure'(beta*v, t) * dt (closures.f90:131-153).  PY is not
    implemented in the reference (returns huge); we raise instead."""
    bv = beta * np.asarray(v, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    x = -bv + t
    if method == HNC:
        with np.errstate(over="ignore"):
            return np.expm1(x) * dt
    if method == KH:
        return np.where(x <= 0.0, np.expm1(np.minimum(x, 0.0)) * dt, 0.0)
    if PSE0 <= method <= PSE7:
        return pse(order(method) - 1, x) * dt
    raise ValueError(f"closure1 not implemented for {method}")


def form(method, x, h, c, cl):
    """Chemical-potential density (closures.f90:574-612):
    HNC: -c - h(c+cl)/2 + h^2/2; KH: drop h^2/2 where h >= 0;
    PY (GF): -c - h(c+cl)/2; PSEn: subtract x^(n+1)/(n+1)! where h >= 0."""
    base = -c - h * (c + cl) / 2.0
    if method == HNC:
        return base + h * h / 2.0
    if method == KH:
        return np.where(h < 0.0, base + h * h / 2.0, base)
    if method == PY:
        return base
    if PSE0 <= method <= PSE7:
        m = order(method) + 1
        mx = float(math.factorial(m))
        return np.where(h < 0.0, base + h * h / 2.0,
                        base + (h * h / 2.0 - x ** m / mx))
    raise ValueError(f"no such closure {method}")


def form1(method, x, dx, h, dh, c, dc, cl):
    """Differential of form() (closures.f90:615-653)."""
    base = -dc - dh * (c + cl) / 2.0 - h * dc / 2.0
    if method == HNC:
        return base + h * dh
    if method == KH:
        return np.where(h < 0.0, base + h * dh, base)
    if method == PY:
        return base
    if PSE0 <= method <= PSE7:
        n = order(method)
        nx = float(math.factorial(n))
        return np.where(h < 0.0, base + h * dh,
                        base + (h * dh - x ** n * dx / nx))
    raise ValueError(f"no such closure {method}")
