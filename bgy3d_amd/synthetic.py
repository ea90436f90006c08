"""Seeded synthetic solutes for the batched 3D-RISM workload.

A solute is generated from its seed with numpy's PCG64 (``default_rng``):

* n ~ U{5..30} sites;
* positions: a random-walk chain, bond length U[1.0, 1.5] A in a uniformly
  random direction, a step re-drawn while any non-bonded pair would be
  closer than 2.0 A; the chain is centred on its centroid (r = 0 is the grid
  centre, as bgy3d-vec.h:597 places it);
* sigma ~ U[2.5, 4.0] A, epsilon ~ U[0.05, 0.25] kcal/mol;
* charges ~ N(0, 0.4^2), shifted uniformly so the net charge is
  Q = seed mod 3 - 1 (0 or +-1).

``box_for`` picks the cubic box: edge = extent + 2 * 8 A padding, and the
smallest N in {64, 96, 128} with spacing in [0.25, 0.35] A.
"""
import math

import numpy as np


def solute(seed, n=None, net=None):
    """Sites of the solute of this seed; n and the net charge default to the
    draws described above."""
    rng = np.random.default_rng(seed)
    if n is None:
        n = int(rng.integers(5, 31))
    x = np.zeros((n, 3))
    for i in range(1, n):
        for _attempt in range(10000):
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            p = x[i - 1] + rng.uniform(1.0, 1.5) * d
            if i < 2 or np.min(np.linalg.norm(x[:i - 1] - p, axis=1)) >= 2.0:
                break
        else:
            raise RuntimeError(f"seed {seed}: could not place site {i}")
        x[i] = p
    x -= x.mean(axis=0)
    sigma = rng.uniform(2.5, 4.0, n)
    eps = rng.uniform(0.05, 0.25, n)
    q = rng.normal(0.0, 0.4, n)
    Q = seed % 3 - 1 if net is None else net
    q += (Q - q.sum()) / n
    return tuple((f"S{i}", tuple(map(float, x[i])), float(sigma[i]), float(eps[i]), float(q[i]))
                 for i in range(n))


def extent(sites):
    x = np.array([s[1] for s in sites])
    return float(np.max(x.max(axis=0) - x.min(axis=0)))


def box_for(sites, pad=8.0):
    """(N, edge) with edge >= extent + 2 pad and 0.25 <= edge/N <= 0.35."""
    edge = extent(sites) + 2 * pad
    for N in (64, 96, 128):
        if edge <= 0.35 * N:
            return N, max(edge, 0.25 * N)
    raise ValueError("solute too large for a 128^3 grid at <= 0.35 A spacing")


def net_charge(sites):
    return sum(s[4] for s in sites)


def batch(seeds):
    return [solute(s) for s in seeds]


def forced_box(sites_list, N=128, pad=8.0):
    """One cubic box for a batch on an N^3 grid: the largest extent + 2 pad,
    at least 0.25 N (spacing >= 0.25 A)."""
    edge = max(extent(s) for s in sites_list) + 2 * pad
    edge = max(edge, 0.25 * N)
    if edge > 0.35 * N:
        raise ValueError("batch does not fit at <= 0.35 A spacing")
    return math.ceil(edge * 4) / 4
