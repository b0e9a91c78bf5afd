"""Run loop around a registered solver (simulation_controller.py:163-1181, reduced).

``SimulationController(cfg, members=None)`` builds the grid, pump, reservoir,
noise, state and model, creates ``cfg.solver.method`` through the registry and
runs ``advance`` in chunks.  Supported run-loop features: total time,
steady-state stop per member (simulation_controller.py:710-734, evaluated on
the device by ``pgpe_density_delta``), in-run accumulators, SIGTERM/SIGUSR1 stop
requests that end the run after the current chunk (:1045-1051), and checkpoints:
periodic (``checkpoint.every``), at a stop request and at the end of a
checkpointed run, and ``checkpoint.resume_from`` (:736-923).  A checkpoint holds
the version-2 content of the reference's (field, field_k, reservoir fields, time
and step, configuration, stop times, steady-state reference, adaptive controller,
accumulator sums) in a NumPy ``.npz`` container (h5py is not a dependency here),
written atomically; resuming continues the run bit for bit.  Storage and
visualisation are outside this package.
"""
from __future__ import annotations

import copy
import json
import math
import os
import signal
import tempfile
import threading
from dataclasses import asdict
from pathlib import Path

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


def _read_checkpoint(path):
    """(attrs, arrays) of a checkpoint written by save_checkpoint."""
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files if k != "attrs"}
        attrs = json.loads(str(z["attrs"]))
    return attrs, arrays


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
        self._seed_drawn = cfg.noise.seed is None and cfg.checkpoint.resume_from is None
        if cfg.noise.seed is None:
            if cfg.checkpoint.resume_from is not None:
                cfg.noise.seed = int(_read_checkpoint(cfg.checkpoint.resume_from)[0]["config"]["noise"]["seed"])
            else:
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
        self._resumed = False
        if cfg.checkpoint.resume_from is not None:
            self.load_checkpoint(cfg.checkpoint.resume_from)
        # result frames (simulation_controller.py:331-332, 395-420): NPY / JSON backends
        from .results import FrameRecorder, StorageContext, storage_names
        self.results = None
        self.run_directory = None
        if cfg.result.save_results and storage_names(cfg.result):
            members = self.state.batch_size if self._member_cfgs is not None else None
            context = StorageContext(
                members=members,
                member_indices=tuple(range(members)) if members is not None else None,
                member_seeds=tuple(int(m.noise.seed) for m in self._member_cfgs) if members is not None else None)
            self.results = FrameRecorder(self, context=context)
            self.run_directory = self.results.output_dir

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
        if self.state.step > n_steps:
            raise ValueError("checkpoint step exceeds configured total_time")
        ck = self.cfg.checkpoint
        interval = _steps(ck.every, dt, "checkpoint.every") if ck.every else 0
        checkpointing = interval > 0 or ck.path is not None or ck.resume_from is not None
        last_checkpoint = -1
        steady = _steps(self.cfg.stop.check_every, dt, "stop.check_every") if self.cfg.stop.steady_state else 0
        previous = {}

        def request_stop(signum, frame):
            self._stop_requested = True

        if self.cfg.checkpoint.on_signal and threading.current_thread() is threading.main_thread():
            for signum in (signal.SIGTERM, signal.SIGUSR1):
                previous[signum] = signal.getsignal(signum)
                signal.signal(signum, request_stop)
        schedules = self._accumulator_schedules(dt, n_steps)
        recording = self.results is not None
        save_steps = (_steps(self.cfg.result.save_every, dt, "result.save_every")
                      if recording and self.cfg.result.save_every is not None else 1)
        try:
            step = self.state.step
            if step == 0 and not self._resumed:
                for acc, start, _, stop in schedules:
                    if start == 0 and stop >= 0 and acc.samples == 0:
                        acc.update(self.state, self.state.t)
                if recording:
                    self.results.record(self.state.t)
            while step < n_steps and not self._stop_requested and not (steady and np.isfinite(self.stop_times).all()):
                end = n_steps
                if steady:
                    end = min(end, (step // steady + 1) * steady)
                if previous:
                    end = min(end, step + 1000)
                if interval:
                    end = min(end, (step // interval + 1) * interval)
                if recording:
                    end = min(end, (step // save_steps + 1) * save_steps)
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
                stopped = bool(steady) and bool(np.isfinite(self.stop_times).all())
                if recording and (step % save_steps == 0 or stopped):
                    self.results.record(self.state.t)
                if interval and step % interval == 0:
                    self.save_checkpoint(self._checkpoint_path())
                    last_checkpoint = step
            if (checkpointing or self._stop_requested) and last_checkpoint != self.state.step:
                self.save_checkpoint(self._checkpoint_path())
        finally:
            for signum, handler in previous.items():
                signal.signal(signum, handler)
            if self._accumulators:
                finalized = [(acc, result) for acc in self._accumulators.values() if acc.samples
                             for result in acc.results()]
                self.accumulated = {result.name: result for _, result in finalized}
            if recording:
                try:
                    if self._accumulators and self.accumulated:
                        self.results.set_accumulated(self.accumulated, {
                            result.name: {"samples": acc.samples, "t_first": acc.t_first, "t_last": acc.t_last}
                            for acc, result in finalized})
                    if steady:
                        self.results.set_stop_times(self.stop_times)
                finally:
                    self.results.finalize()

    # -- checkpoints ---------------------------------------------------------------
    def _checkpoint_path(self) -> Path:
        """checkpoint.path, else the run directory of the result frames, else
        result.output_directory, holding checkpoint.npz."""
        if self.cfg.checkpoint.path is not None:
            return Path(self.cfg.checkpoint.path)
        if self.run_directory is not None:
            return Path(self.run_directory) / "checkpoint.npz"
        return Path(self.cfg.result.output_directory) / "checkpoint.npz"

    def _member_json(self):
        if self._member_cfgs is None:
            return None
        return json.dumps([{name: asdict(m)[name] for name in ("physics", "laser", "potential", "drive", "noise")}
                           for m in self._member_cfgs], sort_keys=True, default=str)

    def save_checkpoint(self, path) -> None:
        """Atomically write the run state (simulation_controller.py:736-798, npz container)."""
        target = Path(path)
        target.parent.mkdir(parents=True, exist_ok=True)
        st = self.state
        arrays = {"field": st.field.get()}
        if st.field_k is not None:
            arrays["field_k"] = st.field_k.get()
        for name, value in zip(self.reservoir.state_names, self.reservoir.get_state()):
            arrays[f"reservoir/{name}"] = value.get()
        attrs = dict(checkpoint_version=2, step=st.step, t=st.t, t0=st.t0, dt=self.cfg.solver.dt, ndim=2,
                     shape=list(self.grid.shape), components=st.components, batched=st.batched,
                     reservoir_type=self.cfg.reservoir.reservoir_type,
                     config=json.loads(json.dumps(asdict(self.cfg), default=str)),
                     stop_time=[None if not np.isfinite(v) else float(v) for v in self.stop_times],
                     steady_last_step=self._steady_last_step, member_configs_json=self._member_json(),
                     accumulators={})
        if self.cfg.solver.adaptive:
            ctl = getattr(self.solver, "step_controller", None)
            attrs["step_controller"] = ctl.state_dict() if ctl is not None else None
        if self._steady_norm is not None:
            attrs["steady_norm"] = [float(v) for v in self._steady_norm]
            arrays["steady_density"] = self._steady_old.get().reshape(st.batch_size, *self.grid.shape)
        for name, acc in self._accumulators.items():
            values = acc.state_dict()
            attrs["accumulators"][name] = {k: values[k] for k in ("samples", "t_first", "t_last")}
            for k, v in values.items():
                if k not in ("samples", "t_first", "t_last"):
                    arrays[f"accumulators/{name}/{k}"] = np.asarray(v)
        arrays["attrs"] = np.asarray(json.dumps(attrs))
        fd, tmp = tempfile.mkstemp(prefix=f".{target.name}.", suffix=".tmp", dir=target.parent)
        try:
            with os.fdopen(fd, "wb") as fh:
                np.savez(fh, **arrays)
            os.replace(tmp, target)
        finally:
            if os.path.exists(tmp):
                os.unlink(tmp)

    def load_checkpoint(self, path) -> None:
        """Validate and restore a checkpoint for this configuration (simulation_controller.py:800-923)."""
        attrs, arrays = _read_checkpoint(path)
        cfg = self.cfg
        if attrs.get("checkpoint_version") != 2:
            raise ValueError(f"checkpoint format v{attrs.get('checkpoint_version')} is not supported; supported: 2")
        if attrs["reservoir_type"] != cfg.reservoir.reservoir_type:
            raise ValueError("reservoir type mismatch")
        if attrs["dt"] != cfg.solver.dt:
            raise ValueError("checkpoint dt mismatch")
        saved = attrs["config"]
        if bool(saved["solver"].get("adaptive", False)) != cfg.solver.adaptive or \
                ("step_controller" in attrs) != cfg.solver.adaptive:
            raise ValueError("checkpoint adaptive mode mismatch")
        if saved.get("stop") != json.loads(json.dumps(asdict(cfg.stop), default=str)):
            raise ValueError("checkpoint stop config mismatch")
        if saved.get("result", {}).get("accumulators", {}) != json.loads(json.dumps(cfg.result.accumulators)):
            raise ValueError("checkpoint accumulator config mismatch")
        if set(attrs.get("accumulators", {})) != set(self._accumulators):
            raise ValueError("checkpoint accumulators mismatch")
        if self._seed_drawn:
            drawn = cfg.noise.seed
            cfg.noise.seed = saved["noise"]["seed"]
            for m in self._member_cfgs or []:
                if m.noise.seed == drawn:
                    m.noise.seed = saved["noise"]["seed"]
            self._seed_drawn = False
        if self._member_cfgs is None and any(saved["noise"][k] != getattr(cfg.noise, k)
                                             for k in ("enabled", "seed", "stochastic", "stochastic_strength")):
            raise ValueError("checkpoint noise config mismatch")
        if self._member_cfgs is not None and attrs.get("member_configs_json") != self._member_json():
            raise ValueError("checkpoint member physics, pump, potential or drive mismatch")
        field = arrays["field"]
        st = self.state
        if int(attrs["ndim"]) != 2 or tuple(attrs["shape"]) != tuple(self.grid.shape):
            raise ValueError("checkpoint grid shape mismatch")
        if int(attrs["components"]) != st.components or field.shape[1] != st.components:
            raise ValueError("checkpoint components mismatch")
        if field.shape != st._canonical_shape or bool(attrs["batched"]) != st.batched:
            raise ValueError("checkpoint batch shape mismatch")
        names = self.reservoir.state_names
        res = [arrays.get(f"reservoir/{n}") for n in names]
        if any(r is None or r.shape != self.reservoir.shape for r in res) or \
                {k for k in arrays if k.startswith("reservoir/")} != {f"reservoir/{n}" for n in names}:
            raise ValueError("checkpoint reservoir names or shapes mismatch")
        step, t, t0 = int(attrs["step"]), float(attrs["t"]), float(attrs["t0"])
        if step < 0 or not np.isfinite(t) or not np.isfinite(t0) or t != t0 + step * cfg.solver.dt:
            raise ValueError("checkpoint time and step mismatch")
        stop_times = np.asarray([np.nan if v is None else v for v in attrs["stop_time"]], dtype=float)
        if stop_times.shape != (st.batch_size,):
            raise ValueError("checkpoint stop_time shape mismatch")
        steady_last = int(attrs["steady_last_step"])
        norm = attrs.get("steady_norm")
        density = arrays.get("steady_density")
        if (norm is None) != (density is None):
            raise ValueError("checkpoint steady-state reference is incomplete")
        if steady_last > step or steady_last < -1:
            raise ValueError("checkpoint steady-state check step mismatch")
        field_k = arrays.get("field_k")
        if field_k is not None and field_k.shape != field.shape:
            raise ValueError("checkpoint spectral batch shape mismatch")
        saved_precision = saved["solver"]["precision"]
        cdt = np.complex64 if cfg.solver.precision == "single" else np.complex128
        if saved_precision != cfg.solver.precision:
            import warnings
            warnings.warn("checkpoint precision differs; casting state", UserWarning, stacklevel=2)
        st.field = DeviceArray.from_numpy(field.astype(cdt))
        if field_k is not None:
            st.field_k = DeviceArray.from_numpy(field_k.astype(cdt))
        st.t, st.t0, st.step = t, t0, step
        self.reservoir.set_state(tuple(r for r in res))
        self.stop_times = stop_times
        self._steady_last_step = steady_last
        if norm is not None:
            self._steady_norm = np.asarray(norm, dtype=float)
            r = np.float32 if cfg.solver.precision == "single" else np.float64
            self._steady_buffers().upload(np.asarray(density, dtype=r).reshape(-1))
        else:
            self._steady_norm = None
        for name, acc in self._accumulators.items():
            values = dict(attrs["accumulators"][name])
            prefix = f"accumulators/{name}/"
            values.update({k[len(prefix):]: v for k, v in arrays.items() if k.startswith(prefix)})
            acc.load_state_dict(values)
        self._resumed = True
        self.solver = create_solver(cfg, self.grid, self.model)
        if cfg.solver.adaptive and attrs.get("step_controller") is not None:
            from .adaptive import PIControl
            self.solver.step_controller = PIControl.for_solver(self.solver)
            self.solver.step_controller.load_state_dict(attrs["step_controller"])

    def _accumulator_schedules(self, dt, n_steps):
        """(accumulator, start, every, stop) in steps (simulation_controller.py:1015-1022)."""
        out = []
        for acc in self._accumulators.values():
            start = 0 if acc.start == 0 else _steps(acc.start, dt, "accumulator start")
            every = _steps(acc.every or self.cfg.result.save_every or dt, dt, "accumulator every")
            stop = n_steps if acc.stop is None else (0 if acc.stop == 0 else _steps(acc.stop, dt, "accumulator stop"))
            out.append((acc, start, every, min(stop, n_steps)))
        return out
