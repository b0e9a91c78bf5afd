"""Shared plumbing of the HIP solvers: member tables, pumps, buffers, graphs."""
from __future__ import annotations

import numpy as np

from .. import _native as N
from ..device import DeviceArray, Graph, Stream, init
from ..layout import MemberSpec, MemberTable
from ..solver_api import AbstractSolver


def real_dtype(precision):
    return np.float32 if precision == "single" else np.float64


def complex_dtype(precision):
    return np.complex64 if precision == "single" else np.complex128


class _MemberPhysics:
    """Attribute view of one member's physics from a BatchPhysics column set."""

    def __init__(self, batch_physics, m):
        self._bp = batch_physics
        self._m = m

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return self._bp.member(name, self._m)


class HipSolverBase(AbstractSolver):
    """Builds the packed member table and the device pump of a uniform batch."""

    def __init__(self, config, grid, model):
        super().__init__(config, grid, model)
        init(getattr(config.compute_engine, "gpu_device", 0))
        self.grid = grid
        self.precision = config.solver.precision
        self.prec_code = 0 if self.precision == "single" else 1
        self.rdtype = real_dtype(self.precision)
        self.cdtype = complex_dtype(self.precision)
        self.reservoir_kind = model.reservoir.kind
        self.nres = N.RESERVOIR_FIELDS[self.reservoir_kind]
        self.components = model.components
        self.stream = Stream()
        self.tile_h = 16
        self._graphs: dict = {}
        self._table = None

    # -- member table -------------------------------------------------------
    def member_specs(self, model, batch: int):
        configs = model.member_configs or [self.config] * batch
        noise = model.noise
        specs = []
        for m in range(batch):
            cfg = configs[m] if m < len(configs) else configs[0]
            strength = float(noise.strength[m]) if (noise is not None and noise.active) else 0.0
            seed = int(noise.seeds[m]) if noise is not None else int(cfg.noise.seed or 0)
            specs.append(MemberSpec(
                nx=self.grid.nx, ny=self.grid.ny, dx=self.grid.dx, dy=self.grid.dy,
                physics=_MemberPhysics(self.batch_physics, m), reservoir=self.reservoir_kind,
                components=self.components, res_components=getattr(model.reservoir, "components", 1),
                pump_off=0, seed=seed, noise_strength=strength, dt=float(self.config.solver.dt),
                circular_polarization=float(cfg.laser.circular_polarization)))
        return specs

    def ensure_table(self, model, batch: int, tile_h: int | None = None) -> MemberTable:
        tile_h = tile_h or self.tile_h
        if self._table is None or self._table.members != batch or self._table.tile_h != tile_h:
            specs = self.member_specs(model, batch)
            envs = self.pump_envelopes(model, batch)
            per = self.grid.nx * self.grid.ny
            for m, s in enumerate(specs):
                s.pump_off = 0 if envs.shape[0] == 1 else m * per
            self._table = MemberTable(specs, tile_h)
            self._pump_dev = DeviceArray.from_numpy(envs.astype(self.rdtype))
            self._graphs.clear()
        return self._table

    # -- pump -----------------------------------------------------------------
    def pump_envelopes(self, model, batch):
        pump = model.pump
        if hasattr(pump, "envelopes"):
            return np.asarray(pump.envelopes())
        terms = pump.separable_terms()
        if len(terms) != 1:
            raise ValueError("the HIP solvers take exactly one separable pump term")
        return np.asarray(terms[0][0])[None]

    def pump_coefficients(self, model, times, batch):
        """(len(times), batch) float64 coefficients of the member pumps."""
        pump = model.pump
        if hasattr(pump, "coefficients"):
            return np.stack([pump.coefficients(t) for t in times])
        (_, coef), = pump.separable_terms()
        return np.asarray([[coef(t)] * batch for t in times], dtype=np.float64)

    def pump_time_independent(self, model) -> bool:
        pump = model.pump
        lasers = [ls[0] for ls in pump.member_lasers] if hasattr(pump, "member_lasers") else pump.lasers
        return all(getattr(l, "time_independent", False) for l in lasers)

    # -- graph helper ---------------------------------------------------------
    def capture(self, key, body) -> Graph:
        g = self._graphs.get(key)
        if g is None:
            self.stream.synchronize()
            g = Graph.capture(self.stream, body)
            if len(self._graphs) >= 4:
                self._graphs.pop(next(iter(self._graphs)))
            self._graphs[key] = g
        return g
