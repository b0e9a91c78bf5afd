"""Pump lasers as cached spatial envelopes times scalar time coefficients.

A laser is ``P(x, y, t) = envelope(x, y) * coefficient(t)``
(laser/abstract_laser.py:74-99).  Supported lasers:

* ``continuous-gaussian`` -- laser/continuous_gaussian.py:14-45
* ``pulse-gaussian`` (peak-amplitude definition) -- laser/pulse_gaussian.py:19-133
* ``profile`` with a constant envelope and the ``gaussian``, ``supergaussian``, ``ring``,
  ``lattice`` (square, triangular, honeycomb, kagome, lieb; Gaussian or super-Gaussian wells,
  disorder, defects, rotation) or ``uniform`` profiles -- laser/profile_laser.py:26-50,
  profiles/builtin_profiles.py:28-240

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
# Spatial profiles of profile lasers (profiles/builtin_profiles.py:28-240).  Each takes the
# grid and the profile mapping (with x0 / y0 filled in) and returns a real (ny, nx) array.
def _r2(grid, params):
    return (grid.X - params.get("x0", 0.0)) ** 2 + (grid.Y - params.get("y0", 0.0)) ** 2


def _positive(params, key):
    v = params.get(key)
    if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(v) or v <= 0:
        raise ValueError(f"profile.{key} must be a positive finite number")
    return float(v)


def profile_gaussian(grid, params):
    return np.exp(-0.5 * _r2(grid, params) / _positive(params, "sigma") ** 2)


def profile_supergaussian(grid, params):
    sigma, order = _positive(params, "sigma"), _positive(params, "order")
    return np.exp(-0.5 * (_r2(grid, params) / sigma ** 2) ** order)


def profile_ring(grid, params):
    radius, width = _positive(params, "radius"), _positive(params, "width")
    return np.exp(-0.5 * ((np.sqrt(_r2(grid, params)) - radius) / width) ** 2)


def profile_uniform(grid, params):
    return np.ones(grid.shape)


# Bravais lattices: primitive vectors and sublattice offsets in units of the lattice constant a
# (a is the nearest-neighbour distance except for honeycomb, where it is sqrt(3) times it).
_R3 = math.sqrt(3.0)
LATTICES = {
    "square": (((1.0, 0.0), (0.0, 1.0)), ((0.0, 0.0),)),
    "triangular": (((1.0, 0.0), (0.5, _R3 / 2)), ((0.0, 0.0),)),
    "honeycomb": (((1.0, 0.0), (0.5, _R3 / 2)), ((0.0, 0.0), (0.5, 0.5 / _R3))),
    "kagome": (((2.0, 0.0), (1.0, _R3)), ((0.0, 0.0), (1.0, 0.0), (0.5, _R3 / 2))),
    "lieb": (((2.0, 0.0), (0.0, 2.0)), ((0.0, 0.0), (1.0, 0.0), (0.0, 1.0))),
}


def _lattice_params(params):
    kind = params.get("lattice")
    if kind not in LATTICES:
        raise ValueError(f"profile.lattice must be one of {sorted(LATTICES)}")
    a, width = _positive(params, "constant"), _positive(params, "width")
    well = params.get("well", "gaussian")
    if well not in ("gaussian", "supergaussian"):
        raise ValueError("profile.well must be gaussian or supergaussian")
    order = _positive(params, "order") if well == "supergaussian" else 1.0
    if well == "gaussian" and "order" in params:
        raise ValueError("profile.order requires well: supergaussian")
    disorder = params.get("disorder", 0.0)
    if isinstance(disorder, bool) or not isinstance(disorder, (int, float)) or not math.isfinite(disorder) \
            or disorder < 0:
        raise ValueError("profile.disorder must be a nonnegative finite number")
    seed = params.get("seed", 0)
    if type(seed) is not int or seed < 0:
        raise ValueError("profile.seed must be a nonnegative integer")
    defects = params.get("defects", [])
    subl = len(LATTICES[kind][1])
    if not isinstance(defects, list) or any(
            not isinstance(d, (list, tuple)) or len(d) != 3 or any(type(v) is not int for v in d)
            or not 0 <= d[2] < subl for d in defects):
        raise ValueError("profile.defects must be a list of [i, j, s] integer indices with a valid sublattice")
    return kind, a, width, order, float(disorder), seed, {tuple(d) for d in defects}


def lattice_sites(lx, ly, params):
    """(i, j, s, x, y) of the lattice sites that cover the box plus one primitive cell, in the
    order i, then j, then sublattice s (profiles/builtin_profiles.py:200-224): the cell index
    range spans the box corners mapped into cell coordinates, widened by two cells; a site is kept
    when both its coordinates are within half the box plus the longer primitive vector."""
    kind, a, _, _, _, _, defects = _lattice_params(params)
    vec, offs = LATTICES[kind]
    th = math.radians(float(params.get("rotation", 0.0)))
    rot = np.array([[math.cos(th), -math.sin(th)], [math.sin(th), math.cos(th)]])
    basis = (a * np.asarray(vec)) @ rot.T
    offsets = (a * np.asarray(offs)) @ rot.T
    origin = np.array([params.get("x0", 0.0), params.get("y0", 0.0)])
    corners = np.array([[sx * lx / 2, sy * ly / 2] for sx in (-1, 1) for sy in (-1, 1)])
    cell = (corners - origin) @ np.linalg.inv(basis)
    lo = np.floor(cell.min(axis=0)).astype(int) - 2
    hi = np.ceil(cell.max(axis=0)).astype(int) + 2
    reach = max(np.linalg.norm(basis[0]), np.linalg.norm(basis[1]))
    out = []
    for i in range(lo[0], hi[0] + 1):
        for j in range(lo[1], hi[1] + 1):
            for k, off in enumerate(offsets):
                if (i, j, k) in defects:
                    continue
                x, y = origin + i * basis[0] + j * basis[1] + off
                if abs(x) <= lx / 2 + reach and abs(y) <= ly / 2 + reach:
                    out.append((i, j, k, float(x), float(y)))
    return out


def profile_lattice(grid, params):
    """Sum of wells exp(-0.5 (r^2 / w^2)^order) on the lattice sites, each of depth 1, or
    1 + N(0, disorder) drawn in site order from default_rng(seed) when disorder > 0."""
    _, _, width, order, disorder, seed, _ = _lattice_params(params)
    rng = np.random.default_rng(seed)
    out = np.zeros(grid.shape, dtype=np.float64)
    for *_, x, y in lattice_sites(grid.lx, grid.ly, params):
        depth = 1.0 + rng.normal(0.0, disorder) if disorder else 1.0
        q = ((grid.X - x) ** 2 + (grid.Y - y) ** 2) / width ** 2
        out += depth * np.exp(-0.5 * (q ** order if order != 1.0 else q))
    return out


PROFILES = {"gaussian": profile_gaussian, "supergaussian": profile_supergaussian, "ring": profile_ring,
            "lattice": profile_lattice, "uniform": profile_uniform}


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
