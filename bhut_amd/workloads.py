"""Synthetic N-body inputs of the benchmark configurations (float64, seeded).

* ``gaussian``: i.i.d. standard normal positions, unit masses (N = 10 000).
* ``plummer``: Plummer sphere, r = a / sqrt(u^(-2/3) - 1) resampled while
  r >= rmax, isotropic directions, equal masses 1/N.
* ``disk_halo``: exponential disk (R ~ Gamma(2, scale), sech^2 vertical
  profile) plus Hernquist halo (r = a sqrt(u) / (1 - sqrt(u)), truncated),
  halo:disk mass 5:1, equal masses within each component.
* ``ensemble``: many independent Plummer spheres of log-uniform size.
"""

from __future__ import annotations

import numpy as np


def gaussian(n: int = 10_000, seed: int = 0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, 3)), np.ones(n)


def _isotropic(rng, r):
    ct = rng.uniform(-1.0, 1.0, len(r))
    ph = rng.uniform(0.0, 2.0 * np.pi, len(r))
    st = np.sqrt(1.0 - ct * ct)
    return np.stack([r * st * np.cos(ph), r * st * np.sin(ph), r * ct], axis=1)


def plummer(n: int, seed: int = 1, a: float = 1.0, rmax_over_a: float = 100.0):
    rng = np.random.default_rng(seed)
    r = np.empty(n)
    fill = 0
    while fill < n:
        u = rng.uniform(0.0, 1.0, n - fill)
        with np.errstate(divide="ignore"):
            rr = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        rr = rr[np.isfinite(rr) & (rr < rmax_over_a * a)]
        r[fill:fill + len(rr)] = rr
        fill += len(rr)
    return _isotropic(rng, r), np.full(n, 1.0 / n)


def plummer_velocities(n: int, seed: int = 4, sigma: float = 0.3):
    rng = np.random.default_rng(seed + 10_000)
    return rng.normal(0.0, sigma, (n, 3))


def disk_halo(n_disk: int = 8_000_000, n_halo: int = 8_000_000, seed: int = 3):
    rng = np.random.default_rng(seed)
    R = rng.gamma(2.0, 3.0, n_disk)
    z = 0.3 * np.arctanh(rng.uniform(-1.0, 1.0, n_disk) * (1 - 1e-12))
    phi = rng.uniform(0.0, 2.0 * np.pi, n_disk)
    disk = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    r = np.empty(n_halo)
    fill = 0
    while fill < n_halo:
        s = np.sqrt(rng.uniform(0.0, 1.0, n_halo - fill))
        with np.errstate(divide="ignore"):
            rr = 10.0 * s / (1.0 - s)
        rr = rr[rr < 200.0]
        r[fill:fill + len(rr)] = rr
        fill += len(rr)
    halo = _isotropic(rng, r)
    pos = np.concatenate([disk, halo])
    mass = np.concatenate([np.full(n_disk, (1.0 / 6.0) / n_disk), np.full(n_halo, (5.0 / 6.0) / n_halo)])
    return pos, mass


def ensemble(n_systems: int = 512, nmin: int = 2_000, nmax: int = 200_000, seed0: int = 1000):
    """List of (positions, masses) for independent Plummer systems."""
    out = []
    for i in range(n_systems):
        rng = np.random.default_rng(seed0 + i)
        a = rng.uniform(0.5, 2.0)
        n = int(np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        x, m = plummer(n, seed=seed0 + i, a=a)
        out.append((x, m))
    return out
