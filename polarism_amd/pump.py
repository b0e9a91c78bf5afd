"""Pump lasers as cached spatial envelopes times scalar time coefficients.

A laser is ``P(x, y, t) = envelope(x, y) * coefficient(t)``
(laser/abstract_laser.py:74-99).  Supported lasers:

* ``continuous-gaussian`` -- laser/continuous_gaussian.py:14-45
* ``pulse-gaussian`` (peak-amplitude definition) -- laser/pulse_gaussian.py:19-133
* ``profile`` with a constant envelope and the ``gaussian``, ``ring``,
  ``lattice`` (square/triangular) or ``uniform`` profiles --
  laser/profile_laser.py:26-50, profiles/builtin_profiles.py:38-74, 176-240

``PumpField.separable_terms()`` exposes ``[(envelope, coefficient_fn)]`` as the
reference does (laser/pump_field.py:37-89); the HIP solvers upload the
envelopes once and the per-stage coefficients per chunk.
"""
from __future__ import annotations

import math

import numpy as np


def _real(precision):
    return np.float32 if precision == "single" else np.float64


# ------------------------------------------------------------------ profiles
def profile_gaussian(grid, params):
    r2 = (grid.X - params.get("x0", 0.0)) ** 2 + (grid.Y - params.get("y0", 0.0)) ** 2
    return np.exp(-0.5 * r2 / params["sigma"] ** 2)


def profile_ring(grid, params):
    r = np.sqrt((grid.X - params.get("x0", 0.0)) ** 2 + (grid.Y - params.get("y0", 0.0)) ** 2)
    return np.exp(-0.5 * ((r - params["radius"]) / params["width"]) ** 2)


def profile_uniform(grid, params):
    return np.ones(grid.shape)


def _lattice_basis(kind, a):
    if kind == "square":
        return np.array([[a, 0.0], [0.0, a]])
    if kind == "triangular":
        return np.array([[a, 0.0], [a / 2, math.sqrt(3) * a / 2]])
    raise ValueError(f"unsupported lattice {kind!r} (square or triangular)")


def profile_lattice(grid, params):
    """Sum of unit Gaussian wells on a Bravais lattice covering the box plus a margin."""
    basis = _lattice_basis(params["lattice"], float(params["constant"]))
    th = math.radians(params.get("rotation", 0.0))
    rot = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    basis = basis @ rot.T
    origin = np.array([params.get("x0", 0.0), params.get("y0", 0.0)])
    corners = np.array([[x, y] for x in (-grid.lx / 2, grid.lx / 2) for y in (-grid.ly / 2, grid.ly / 2)])
    cells = (corners - origin) @ np.linalg.inv(basis)
    lo = np.floor(cells.min(axis=0)).astype(int) - 2
    hi = np.ceil(cells.max(axis=0)).astype(int) + 2
    margin = max(np.linalg.norm(basis[0]), np.linalg.norm(basis[1]))
    width = float(params["width"])
    disorder = float(params.get("disorder", 0.0))
    rng = np.random.default_rng(params.get("seed", 0))
    out = np.zeros(grid.shape)
    for i in range(lo[0], hi[0] + 1):
        for j in range(lo[1], hi[1] + 1):
            x, y = origin + i * basis[0] + j * basis[1]
            if abs(x) <= grid.lx / 2 + margin and abs(y) <= grid.ly / 2 + margin:
                depth = 1.0 + rng.normal(0.0, disorder) if disorder else 1.0
                out += depth * np.exp(-0.5 * ((grid.X - x) ** 2 + (grid.Y - y) ** 2) / width ** 2)
    return out


PROFILES = {"gaussian": profile_gaussian, "ring": profile_ring, "lattice": profile_lattice,
            "uniform": profile_uniform}


def create_profile(spec, grid):
    kind = spec.get("type")
    if kind not in PROFILES:
        raise ValueError(f"Unknown profile '{kind}'. Available: {sorted(PROFILES)}")
    return PROFILES[kind](grid, spec)


# ------------------------------------------------------------------ lasers
class Laser:
    """Base class: cached real envelope and a scalar coefficient of time."""

    def __init__(self, cfg, grid, precision):
        self.P0 = float(cfg.P0)
        self.x0 = float(cfg.x0)
        self.y0 = float(cfg.y0)
        self.delay = float(cfg.delay)
        self.precision = precision
        self.envelope = self.space(grid).astype(_real(precision), copy=False)

    def space(self, grid):
        raise NotImplementedError

    def amplitude(self, t):
        return self.P0

    def time_profile(self, t):
        return 1.0

    def power_coefficient(self, t: float) -> float:
        if t < self.delay:
            return 0.0
        te = t - self.delay
        return self.amplitude(te) * self.time_profile(te)

    time_independent = True


class ContinuousGaussian(Laser):
    def __init__(self, cfg, grid, precision):
        self.sigma_space = float(cfg.sigma_space)
        super().__init__(cfg, grid, precision)

    def space(self, grid):
        r2 = (grid.X - self.x0) ** 2 + (grid.Y - self.y0) ** 2
        return np.exp(-0.5 * r2 / self.sigma_space ** 2)


class PulseGaussian(ContinuousGaussian):
    time_independent = False

    def __init__(self, cfg, grid, precision):
        if cfg.power_definition != "peak_amplitude":
            raise ValueError("pulse-gaussian supports power_definition='peak_amplitude' only")
        super().__init__(cfg, grid, precision)
        self.Pmax = float(cfg.Pmax)
        self.sigma_time = float(cfg.sigma_time)
        self.pulse_separation = float(cfg.pulse_separation)
        self.cutoff_sigma = float(cfg.cutoff_sigma)
        self.n_pulses = int(cfg.n_pulses)
        self.phase = self.cutoff_sigma * self.sigma_time

    def _index(self, t):
        return max(0, round((t - self.phase) / self.pulse_separation))

    def amplitude(self, t):
        n = self._index(t)
        if self.n_pulses > 0 and n >= self.n_pulses:
            return 0.0
        return min(self.P0 + n * (self.Pmax - self.P0), self.Pmax)

    def time_profile(self, t):
        n = self._index(t)
        if self.n_pulses > 0 and n >= self.n_pulses:
            return 0.0
        d = t - n * self.pulse_separation - self.phase
        if abs(d) > self.cutoff_sigma * self.sigma_time:
            return 0.0
        return math.exp(-0.5 * (d / self.sigma_time) ** 2)


class ProfileLaser(Laser):
    def __init__(self, cfg, grid, precision):
        if not cfg.profile:
            raise ValueError("laser.profile is required for laser_type='profile'")
        if cfg.envelope.get("type", "constant") != "constant":
            raise ValueError("profile lasers support the constant envelope only")
        self.profile = dict(cfg.profile)
        super().__init__(cfg, grid, precision)

    def space(self, grid):
        spec = dict(self.profile)
        spec.setdefault("x0", self.x0)
        spec.setdefault("y0", self.y0)
        return create_profile(spec, grid)


LASERS = {"continuous-gaussian": ContinuousGaussian, "pulse-gaussian": PulseGaussian,
          "profile": ProfileLaser}


def create_lasers(cfg, grid, precision):
    if cfg.mode != "single":
        raise ValueError("only laser.mode='single' is supported")
    if cfg.laser_type not in LASERS:
        raise ValueError(f"Unknown laser '{cfg.laser_type}'. Available: {sorted(LASERS)}")
    return [LASERS[cfg.laser_type](cfg, grid, precision)]


class PumpField:
    """Sum of separable lasers for one member (laser/pump_field.py:37-89)."""

    def __init__(self, lasers, grid, precision):
        self.lasers = lasers
        self.grid = grid
        self.precision = precision

    def separable_terms(self):
        return [(laser.envelope, laser.power_coefficient) for laser in self.lasers]

    def __call__(self, t):
        total = np.zeros(self.grid.shape, dtype=_real(self.precision))
        for env, coef in self.separable_terms():
            total = total + env * coef(t)
        return total


class BatchedPumpField:
    """Member-stacked pumps (laser/pump_field.py:92-152): one laser slot per member."""

    def __init__(self, member_lasers, grid, precision):
        if len({len(ls) for ls in member_lasers}) != 1 or len(member_lasers[0]) != 1:
            raise ValueError("batched pumps need exactly one laser per member")
        self.member_lasers = member_lasers
        self.grid = grid
        self.precision = precision
        envs = [ls[0].envelope for ls in member_lasers]
        self.shared = all(np.array_equal(envs[0], e) for e in envs[1:])

    def envelopes(self):
        """(members or 1, ny, nx) real envelopes."""
        if self.shared:
            return self.member_lasers[0][0].envelope[None]
        return np.stack([ls[0].envelope for ls in self.member_lasers])

    def coefficients(self, t) -> np.ndarray:
        return np.asarray([ls[0].power_coefficient(t) for ls in self.member_lasers], dtype=np.float64)

    def __call__(self, t):
        envs = self.envelopes()
        coef = self.coefficients(t).astype(_real(self.precision))
        return (envs * coef[:, None, None])[:, None] if not self.shared else envs[None] * coef[:, None, None, None]
