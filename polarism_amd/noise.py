"""Initial condensate field and the truncated-Wigner increment.

* ``make_initial_psi`` -- host NumPy ``default_rng(seed)`` complex Gaussian,
  mean-subtracted and RMS-normalised (noise/initial_field.py:363-380); the
  reference also draws this on the host so CPU and GPU runs start alike.
* ``NoiseApplier`` -- per-member seed, gamma_C, R and strength; adds
  ``sigma (N1 + i N2)`` through ``pgpe_wigner_add`` with Philox counters
  (cell, step, stream = 1 + component) (noise/stochastic.py:15-106).
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .device import DeviceArray
from .layout import MemberSpec, MemberTable


def draw_seed() -> int:
    return int(np.random.SeedSequence().entropy % (1 << 63))


def make_initial_psi(shape, eps, mode="complex_gaussian_zero_mean", seed=None, precision="double"):
    rdtype = np.float32 if precision == "single" else np.float64
    cdtype = np.complex64 if precision == "single" else np.complex128
    if mode == "none" or eps == 0.0:
        return np.zeros(shape, dtype=cdtype)
    rng = np.random.default_rng(seed)
    if mode == "legacy_positive_uniform":
        return (eps * (rng.random(shape, dtype=rdtype) + 1j * rng.random(shape, dtype=rdtype))).astype(cdtype)
    if mode != "complex_gaussian_zero_mean":
        raise ValueError(f"noise.initial={mode!r} is not supported")
    re = rng.standard_normal(shape).astype(rdtype)
    im = rng.standard_normal(shape).astype(rdtype)
    psi = re + 1j * im
    psi -= np.mean(psi)
    rms = float(np.sqrt(np.mean(np.abs(psi) ** 2)))
    if rms < 1e-30:
        rms = 1.0
    return ((eps / rms) * psi).astype(cdtype)


class NoiseApplier:
    """Stateless counter-based Wigner increments for every member of a batch."""

    def __init__(self, config, grid, members=None):
        self.members = members or [config]
        self.precision = config.solver.precision
        self.grid = grid
        self.dt = float(config.solver.dt)
        self.seeds = np.asarray([m.noise.seed for m in self.members], dtype=np.uint64)
        self.gamma = np.asarray([m.physics.gamma_C for m in self.members], dtype=np.float64)
        self.scattering = np.asarray([m.physics.R for m in self.members], dtype=np.float64)
        self.strength = np.asarray([m.noise.stochastic_strength for m in self.members], dtype=np.float64)
        self.active = bool(config.noise.enabled and config.noise.stochastic == "wigner"
                           and np.any(self.strength > 0.0))
        self._table = None
        self._components = None

    def member_specs(self, components=1, reservoir="single"):
        return [MemberSpec(nx=self.grid.nx, ny=self.grid.ny, dx=self.grid.dx, dy=self.grid.dy,
                           physics=m.physics, reservoir=reservoir, components=components,
                           seed=int(self.seeds[i]), noise_strength=float(self.strength[i]), dt=self.dt)
                for i, m in enumerate(self.members)]

    def table(self, components=1) -> MemberTable:
        if self._table is None or self._components != components:
            self._table = MemberTable(self.member_specs(components))
            self._components = components
        return self._table

    def apply_arrays(self, field: DeviceArray, active_density, step: int, stream=None) -> None:
        """Add one increment to a canonical ``(B, C, ny, nx)`` device field."""
        if not self.active:
            return
        if not isinstance(field, DeviceArray):
            raise TypeError("NoiseApplier.apply_arrays needs a DeviceArray field")
        B, Cc = field.shape[:2]
        if B != len(self.members):
            raise ValueError("field batch does not match the noise members")
        tab = self.table(Cc)
        prec = 0 if self.precision == "single" else 1
        active_ptr, active_c = None, 1
        if isinstance(active_density, DeviceArray):
            active_ptr, active_c = active_density.ptr, active_density.shape[1]
        from .device import _sp
        N.call("pgpe_wigner_add", field.ptr, active_ptr, active_c, tab.table_ptr, tab.tile_map_ptr,
               tab.num_tiles, tab.members, prec, int(step), _sp(stream))

    def apply_to_state(self, state, reservoir, step: int) -> None:
        if not self.active:
            return
        active = reservoir.get_active_density(reservoir.get_state())
        self.apply_arrays(state.field, active, step)
        state.invalidate_psi_k()
