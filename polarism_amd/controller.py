"""Run loop around a registered solver (simulation_controller.py:163-1181, reduced).

``SimulationController(cfg, members=None)`` builds the grid, pump, reservoir,
noise, state and model, creates ``cfg.solver.method`` through the registry and
runs ``advance`` in chunks.  Supported run-loop features: total time,
steady-state stop per member (simulation_controller.py:710-734, evaluated on
the device by ``pgpe_density_delta``), and SIGTERM/SIGUSR1 stop requests that
end the run after the current chunk (:1045-1051).  Storage, visualisation and
checkpoints are outside this package.
"""
from __future__ import annotations

import copy
import math
import signal
import threading

import numpy as np

from . import _native as N
from .config import Config
from .device import DeviceArray, init
from .grid import create_grid
from .model import Model
from .noise import NoiseApplier, draw_seed
from .physics import BatchPhysics
from .pump import BatchedPumpField, PumpField, create_lasers
from .reservoir import create_reservoir
from .solver_api import check_solver_compatibility, create_solver
from .state import SimulationState
from .validation import validate_config


def apply_overrides(cfg: Config, overrides: dict) -> Config:
    """Return a copy of ``cfg`` with ``{"section.field": value}`` applied (sweep_config.py:69-124)."""
    out = copy.deepcopy(cfg)
    for key, value in (overrides or {}).items():
        section, _, name = key.partition(".")
        if not name:
            raise ValueError(f"override key {key!r} must be 'section.field'")
        target = getattr(out, section)
        if "." in name:
            head, _, sub = name.partition(".")
            mapping = dict(getattr(target, head))
            mapping[sub] = value
            setattr(target, head, mapping)
        else:
            setattr(target, name, value)
    return out


def _next_sample(step, start, every, stop):
    """First sampling step after ``step`` (simulation_controller.py:1024-1026)."""
    due = start if step < start else start + ((step - start) // every + 1) * every
    return due if due <= stop else None


def _steps(value, dt, what):
    n = round(value / dt)
    if n <= 0 or not math.isclose(n * dt, value, rel_tol=1e-9, abs_tol=1e-12):
        raise ValueError(f"{what}={value} must be a positive multiple of solver.dt={dt}")
    return n


class SimulationController:
    def __init__(self, cfg: Config, members: list[dict] | None = None):
        from . import solvers  # noqa: F401  (registers the HIP methods)
        if members is not None and len(members) == 0:
            raise ValueError("members must contain at least one override mapping")
        cfg = copy.deepcopy(cfg)
        if cfg.noise.seed is None:
            cfg.noise.seed = draw_seed()
        member_cfgs = [apply_overrides(cfg, o) for o in members] if members is not None else None
        for m in member_cfgs or []:
            if m.noise.seed is None:
                m.noise.seed = cfg.noise.seed
        if member_cfgs is not None and len(member_cfgs) == 1:
            cfg = member_cfgs[0]
        batched = member_cfgs is not None and len(member_cfgs) >= 2
        self.cfg = cfg
        self._member_cfgs = member_cfgs if batched else None
        validate_config(cfg)
        check_solver_compatibility(cfg)
        init(cfg.compute_engine.gpu_device)
        precision = cfg.solver.precision
        self.grid = create_grid(cfg.grid)
        components = 2 if cfg.physics.spin == "spinor" else 1
        batch_values = None
        if batched:
            from .config import FLOAT_PHYSICS_FIELDS
            batch_values = {name: np.asarray([getattr(m.physics, name) for m in member_cfgs])
                            for name in sorted(FLOAT_PHYSICS_FIELDS)
                            if any(getattr(m.physics, name) != getattr(member_cfgs[0].physics, name)
                                   for m in member_cfgs[1:])}
            for m in member_cfgs[1:]:
                for section in ("grid", "reservoir", "solver"):
                    if getattr(m, section) != getattr(member_cfgs[0], section):
                        raise ValueError(f"batched members must share the {section} section")
            self.pump = BatchedPumpField([create_lasers(m.laser, self.grid, precision) for m in member_cfgs],
                                         self.grid, precision)
            fields = []
            for m in member_cfgs:
                st = SimulationState(self.grid, m.noise.initial_amplitude if m.noise.enabled else 0.0, precision,
                                     mode=m.noise.initial if m.noise.enabled else "none", seed=m.noise.seed,
                                     components=components)
                fields.append(st.field.get())
            self.state = SimulationState(self.grid, 0.0, precision, batch_size=len(member_cfgs),
                                         components=components,
                                         initial_fields=np.concatenate(fields, axis=0))
        else:
            self.pump = PumpField(create_lasers(cfg.laser, self.grid, precision), self.grid, precision)
            self.state = SimulationState(self.grid, cfg.noise.initial_amplitude if cfg.noise.enabled else 0.0,
                                         precision, mode=cfg.noise.initial if cfg.noise.enabled else "none",
                                         seed=cfg.noise.seed, components=components)
        batch = self.state.batch_size
        self.reservoir = create_reservoir(cfg.reservoir, cfg.physics, self.grid, precision, batch_size=batch)
        physics = BatchPhysics(cfg.physics, batch_values, precision)
        noise = None
        members_all = self._member_cfgs or [cfg]
        if cfg.noise.enabled and cfg.noise.stochastic == "wigner" and any(
                m.noise.stochastic_strength > 0 for m in members_all):
            noise = NoiseApplier(cfg, self.grid, self._member_cfgs)
        pump_weights = None
        if components == 2:
            pump_weights = np.asarray([[(1 + m.laser.circular_polarization) / 2,
                                        (1 - m.laser.circular_polarization) / 2] for m in members_all])
        self.model = Model(self.grid, physics, precision, 0.0, self.pump, self.reservoir, None, noise,
                           pump_weights, getattr(self.reservoir, "components", 1), components,
                           self._member_cfgs)
        self.solver = create_solver(cfg, self.grid, self.model)
        from .accumulators import available_accumulators
        self._accumulators = {name: available_accumulators.lookup(name)(self, **params)
                              for name, params in cfg.result.accumulators.items()}
        self.accumulated: dict = {}
        self.stop_times = np.full(batch, np.nan)
        self._steady_old = None
        self._steady_norm = None
        self._steady_last_step = -1
        self._stop_requested = False
        self.steady_checks = 0

    # -- steady state ------------------------------------------------------------
    def _steady_buffers(self):
        if self._steady_old is None:
            tab = self._steady_table()
            r = np.float32 if self.cfg.solver.precision == "single" else np.float64
            self._steady_old = DeviceArray.zeros((tab.total_points,), r)
            self._steady_partials = DeviceArray.zeros((3 * tab.num_tiles,), np.float64)
            self._steady_sums = DeviceArray.zeros((3 * tab.members,), np.float64)
        return self._steady_old

    def _steady_table(self):
        if not hasattr(self, "_stab"):
            from .layout import MemberSpec, MemberTable
            g = self.grid
            self._stab = MemberTable([MemberSpec(g.nx, g.ny, g.dx, g.dy, self.cfg.physics)
                                      for _ in range(self.state.batch_size)])
        return self._stab

    def density_sums(self, update=True) -> np.ndarray:
        """(B, 3): sum rho, sum (rho - previous)^2, sum previous^2 over each member."""
        tab = self._steady_table()
        old = self._steady_buffers()
        prec = 0 if self.cfg.solver.precision == "single" else 1
        N.call("pgpe_density_delta", self.state.field.ptr, old.ptr, prec, self.state.components,
               tab.table_ptr, tab.tile_map_ptr, tab.num_tiles, tab.members, self._steady_partials.ptr,
               self._steady_sums.ptr, 1 if update else 0, None)
        return self._steady_sums.get().reshape(-1, 3)

    def _check_steady_state(self) -> None:
        sums = self.density_sums(update=True)
        cell = self.grid.dx * self.grid.dy
        norms = sums[:, 0] * cell
        self.steady_checks += 1
        if self._steady_norm is not None:
            rel_d = np.sqrt(sums[:, 1]) / np.maximum(np.sqrt(sums[:, 2]), 1e-30)
            old_norms = self._steady_norm
            rel_n = np.abs(norms - old_norms) / np.maximum(old_norms, 1e-30)
            reached = (norms < 1e-30) | ((rel_n < self.cfg.stop.rtol) & (rel_d < self.cfg.stop.rtol))
            self.last_relative_density = rel_d
            if self.state.t >= self.cfg.stop.min_time:
                self.stop_times[reached & np.isnan(self.stop_times)] = self.state.t
        self._steady_norm = norms.copy()
        self._steady_last_step = self.state.step

    # -- run ------------------------------------------------------------------------
    def run(self) -> None:
        dt = self.cfg.solver.dt
        n_steps = round(self.cfg.solver.total_time / dt)
        steady = _steps(self.cfg.stop.check_every, dt, "stop.check_every") if self.cfg.stop.steady_state else 0
        previous = {}

        def request_stop(signum, frame):
            self._stop_requested = True

        if self.cfg.checkpoint.on_signal and threading.current_thread() is threading.main_thread():
            for signum in (signal.SIGTERM, signal.SIGUSR1):
                previous[signum] = signal.getsignal(signum)
                signal.signal(signum, request_stop)
        schedules = self._accumulator_schedules(dt, n_steps)
        try:
            step = self.state.step
            if step == 0:
                for acc, start, _, stop in schedules:
                    if start == 0 and stop >= 0 and acc.samples == 0:
                        acc.update(self.state, self.state.t)
            while step < n_steps and not self._stop_requested and not (steady and np.isfinite(self.stop_times).all()):
                end = n_steps
                if steady:
                    end = min(end, (step // steady + 1) * steady)
                if previous:
                    end = min(end, step + 1000)
                for _, start, every, stop in schedules:
                    due = _next_sample(step, start, every, stop)
                    if due is not None:
                        end = min(end, due)
                self.solver.advance(self.model, self.state, end - step)
                step = self.state.step
                for acc, start, every, stop in schedules:
                    if start <= step <= stop and (step - start) % every == 0:
                        acc.update(self.state, self.state.t)
                if steady and step % steady == 0 and self._steady_last_step != step:
                    self._check_steady_state()
        finally:
            for signum, handler in previous.items():
                signal.signal(signum, handler)
            if self._accumulators:
                self.accumulated = {result.name: result for acc in self._accumulators.values() if acc.samples
                                    for result in acc.results()}

    def _accumulator_schedules(self, dt, n_steps):
        """(accumulator, start, every, stop) in steps (simulation_controller.py:1015-1022)."""
        out = []
        for acc in self._accumulators.values():
            start = 0 if acc.start == 0 else _steps(acc.start, dt, "accumulator start")
            every = _steps(acc.every or self.cfg.result.save_every or dt, dt, "accumulator every")
            stop = n_steps if acc.stop is None else (0 if acc.stop == 0 else _steps(acc.stop, dt, "accumulator stop"))
            out.append((acc, start, every, min(stop, n_steps)))
        return out
