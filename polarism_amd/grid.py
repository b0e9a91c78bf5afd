"""Simulation grids: periodic (grid/periodic_grid.py:35-69) and closed
interval (grid/closed_interval_grid.py:35-80).

Periodic: ``dx = lx / nx``, centred coordinates ``X = (i - (n-1)/2) dx``.
Closed interval: ``dx = lx / (nx - 1)``, ``X = linspace(-lx/2, lx/2, nx)``
(the end nodes lie on the walls).  Both carry FFT wave numbers
``k = 2 pi fftfreq(n, dx)`` and ``k^2 = KX^2 + KY^2`` on the host in float64;
x is the fastest axis, the field shape is ``(ny, nx)``.
"""
from __future__ import annotations

import numpy as np


class PeriodicGrid:
    boundary = "periodic"
    grid_type = "periodic"

    def __init__(self, cfg):
        if getattr(cfg, "ndim", 2) != 2:
            raise ValueError("the HIP solvers support 2-D grids (grid.ndim = 2)")
        self.ndim = 2
        self.nx, self.ny = int(cfg.nx), int(cfg.ny)
        if self.nx < 1 or self.ny < 1:
            raise ValueError("grid.nx and grid.ny must be positive")
        self.lx, self.ly = float(cfg.lx), float(cfg.ly)
        self.dx = self.lx / self.nx
        self.dy = self.ly / self.ny
        self.shape = (self.ny, self.nx)
        self.spacing = (self.dy, self.dx)
        self.lengths = (self.ly, self.lx)
        x = (np.arange(self.nx, dtype=np.float64) - (self.nx - 1) * 0.5) * self.dx
        y = (np.arange(self.ny, dtype=np.float64) - (self.ny - 1) * 0.5) * self.dy
        self.X, self.Y = np.meshgrid(x, y, indexing="xy")
        self.kx = 2 * np.pi * np.fft.fftfreq(self.nx, d=self.dx)
        self.ky = 2 * np.pi * np.fft.fftfreq(self.ny, d=self.dy)
        self.KX, self.KY = np.meshgrid(self.kx, self.ky, indexing="xy")
        self.k_squared = self.KX ** 2 + self.KY ** 2

    @property
    def cell_volume(self) -> float:
        return self.dx * self.dy


class ClosedIntervalGrid(PeriodicGrid):
    boundary = "closed-interval"
    grid_type = "closed-interval"

    def __init__(self, cfg):
        if getattr(cfg, "ndim", 2) != 2:
            raise ValueError("the HIP solvers support 2-D grids (grid.ndim = 2)")
        self.ndim = 2
        self.nx, self.ny = int(cfg.nx), int(cfg.ny)
        if self.nx < 2 or self.ny < 2:
            raise ValueError("ClosedIntervalGrid requires nx >= 2 and ny >= 2 in 2D")
        self.lx, self.ly = float(cfg.lx), float(cfg.ly)
        self.dx = self.lx / (self.nx - 1)
        self.dy = self.ly / (self.ny - 1)
        self.shape = (self.ny, self.nx)
        self.spacing = (self.dy, self.dx)
        self.lengths = (self.ly, self.lx)
        x = np.linspace(-self.lx / 2, self.lx / 2, self.nx, dtype=np.float64)
        y = np.linspace(-self.ly / 2, self.ly / 2, self.ny, dtype=np.float64)
        self.X, self.Y = np.meshgrid(x, y, indexing="xy")
        self.kx = 2 * np.pi * np.fft.fftfreq(self.nx, d=self.dx)
        self.ky = 2 * np.pi * np.fft.fftfreq(self.ny, d=self.dy)
        self.KX, self.KY = np.meshgrid(self.kx, self.ky, indexing="xy")
        self.k_squared = self.KX ** 2 + self.KY ** 2


def create_grid(cfg) -> PeriodicGrid:
    kind = getattr(cfg, "grid_type", "periodic")
    if kind == "periodic":
        return PeriodicGrid(cfg)
    if kind == "closed-interval":
        return ClosedIntervalGrid(cfg)
    raise ValueError(f"unsupported grid_type {kind!r}: the HIP solvers run periodic or closed-interval grids")
