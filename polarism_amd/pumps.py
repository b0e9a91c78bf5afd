"""Seeded random pump maps for sweeps and surrogate training data.

Pump families (profiles as in profiles/builtin_profiles.py:38-74, 225-240 of
the reference):

* ``gaussians`` -- a mixture of 1-4 Gaussian spots, sigma U[8, 20] um,
  centres within +-20 um of the box centre, weights U[0.5, 1];
* ``ring``      -- a Gaussian ring, radius U[12, 30] um, width U[4, 8] um;
* ``lattice``   -- a square lattice of Gaussian wells, period U[6, 15] um,
  well width period * U[0.3, 0.45], inside a Gaussian window of radius U[25, 40] um.

Every map is normalised to a peak of 1, so the pump is ``P0 * map`` and P0
is the peak pump rate.  Thresholds are those of a uniform pump:
single reservoir ``P_th = gamma_R gamma_C / R``; double reservoir
``P_th = (gamma_C / R) [(gamma_A + R_AI)(gamma_I + R_IA) / R_IA - R_AI]``.
"""
from __future__ import annotations

import numpy as np


def coords(nx, ny, lx, ly):
    x = (np.arange(nx) - (nx - 1) * 0.5) * (lx / nx)
    y = (np.arange(ny) - (ny - 1) * 0.5) * (ly / ny)
    return np.meshgrid(x, y, indexing="xy")


def threshold(physics, reservoir="single") -> float:
    p = physics
    if reservoir == "single":
        return p.gamma_R * p.gamma_C / p.R
    if reservoir == "double":
        return (p.gamma_C / p.R) * ((p.gamma_A + p.R_AI) * (p.gamma_I + p.R_IA) / p.R_IA - p.R_AI)
    raise ValueError(reservoir)


def gaussian_map(X, Y, sigma, x0=0.0, y0=0.0):
    return np.exp(-0.5 * ((X - x0) ** 2 + (Y - y0) ** 2) / sigma ** 2)


def random_pump(rng: np.random.Generator, nx, ny, lx, ly, kind: str | None = None):
    """Return (map with peak 1, description dict)."""
    X, Y = coords(nx, ny, lx, ly)
    kind = kind or rng.choice(["gaussians", "ring", "lattice"])
    half = min(20.0, 0.25 * min(lx, ly))
    if kind == "gaussians":
        k = int(rng.integers(1, 5))
        out = np.zeros_like(X)
        spots = []
        for _ in range(k):
            s, x0, y0, w = rng.uniform(8, 20), rng.uniform(-half, half), rng.uniform(-half, half), rng.uniform(0.5, 1)
            out += w * gaussian_map(X, Y, s, x0, y0)
            spots.append((s, x0, y0, w))
        desc = {"kind": "gaussians", "spots": spots}
    elif kind == "ring":
        r0, w = rng.uniform(12, 30), rng.uniform(4, 8)
        r = np.sqrt(X ** 2 + Y ** 2)
        out = np.exp(-0.5 * ((r - r0) / w) ** 2)
        desc = {"kind": "ring", "radius": r0, "width": w}
    elif kind == "lattice":
        a = rng.uniform(6, 15)
        width = a * rng.uniform(0.3, 0.45)
        window = rng.uniform(25, 40)
        ox, oy = rng.uniform(0, a, size=2)
        # a square lattice of Gaussians is separable: sum_ij g(x - x_i) g(y - y_j)
        n = int(np.ceil(max(lx, ly) / a)) + 2
        cx = ox + a * np.arange(-n, n + 1)
        cy = oy + a * np.arange(-n, n + 1)
        cx = cx[np.abs(cx) <= lx / 2 + a]
        cy = cy[np.abs(cy) <= ly / 2 + a]
        gx = np.exp(-0.5 * (X[0][:, None] - cx[None, :]) ** 2 / width ** 2).sum(axis=1)
        gy = np.exp(-0.5 * (Y[:, 0][:, None] - cy[None, :]) ** 2 / width ** 2).sum(axis=1)
        out = gy[:, None] * gx[None, :]
        out *= gaussian_map(X, Y, window)
        desc = {"kind": "lattice", "period": a, "width": width, "window": window}
    else:
        raise ValueError(kind)
    peak = out.max()
    return out / peak, desc
