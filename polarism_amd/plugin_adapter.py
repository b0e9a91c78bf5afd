"""Wrap the HIP methods as solvers of the reference's own registry.

The reference loads every module under ``./polarism_plugins/`` (polarism/plugins.py:12-41);
``polarism_amd/polarism_plugin.py`` is that module.  It passes the reference's classes in, so
this file imports nothing from the reference and its logic runs (and is tested) without it:

* ``config_mapping(config)``   -- the reference ``Config`` as the nested mapping that
  ``polarism_amd.Config.from_dict`` reads: the same sections and field names
  (config/simulation_parameters.py); keys this package does not declare are left out and
  listed in ``ignored``;
* ``member_overrides(model)``  -- the per-member ``physics.*`` overrides of a batched model
  (``model.physics.host``, config/batch_physics.py:17-64) plus the member noise seeds, or None;
* ``make_hip_solver(name, ...)`` -- an ``AbstractSolver`` subclass for HIP method ``name``
  honouring the reference contract (abstract_solver.py:71-101): ``step`` moves one fixed step
  (``state.step += 1``, ``state.t = t0 + step dt``) or one adaptive trial (``t += h`` when
  accepted, the tick count unchanged); ``advance`` moves ``n`` ticks and lands on
  ``t0 + (step + n) dt``.  The HIP side adds the Wigner increment inside its own kernels, so the
  class declares ``injects_noise = True`` and the reference's loop never adds it a second time.
  The reference state is copied in before and out after each call: the field and reservoir
  arrays the reference holds never alias the HIP device buffers.
"""
from __future__ import annotations

import dataclasses

import numpy as np

HIP_METHODS = ("rk4-fdm-hip", "split-step-fft-hip", "split-step-fft-kouter-hip", "ifrk-fft-hip",
               "ifrk4-fft-hip", "ip-rk4-hip")


def config_mapping(config, ignored: list | None = None) -> dict:
    """Nested {section: {key: value}} of a reference Config restricted to the keys this package
    declares; dropped ``section.key`` paths are appended to ``ignored``."""
    from .config import Config
    data = dataclasses.asdict(config) if dataclasses.is_dataclass(config) else dict(config)
    template = Config()
    out = {}
    for section, values in data.items():
        if section not in Config._declared:
            if ignored is not None:
                ignored.append(section)
            continue
        if section == "terms":
            out[section] = dict(values)
            continue
        known = type(getattr(template, section))._declared
        out[section] = {k: v for k, v in values.items() if k in known}
        if ignored is not None:
            ignored.extend(f"{section}.{k}" for k in values if k not in known)
    return out


def member_overrides(model) -> list | None:
    """One override mapping per member of a batched model, or None for a single member."""
    host = getattr(getattr(model, "physics", None), "host", None) or {}
    size = getattr(model.physics, "size", None)
    if not host or not size:
        return None
    rows = [{f"physics.{name}": float(col[m]) for name, col in host.items()} for m in range(size)]
    seeds = getattr(getattr(model, "noise", None), "seeds", None)
    if seeds is not None and len(seeds) == size:
        for m, row in enumerate(rows):
            row["noise.seed"] = int(seeds[m])
    return rows


def make_hip_solver(name: str, *, base, report_cls, to_host, to_array, controller_factory=None):
    """Build the reference-side solver class for HIP method ``name``.

    ``base`` is the reference ``AbstractSolver``, ``report_cls`` its ``StepReport``;
    ``to_host`` turns a reference array into NumPy and ``to_array`` a NumPy array into the
    reference's array type (compute_engine.to_cpu / xp.asarray).  ``controller_factory(cfg,
    members)`` builds the HIP-side controller (``polarism_amd.SimulationController`` unless
    given)."""
    from .config import Config
    if controller_factory is None:
        from .controller import SimulationController as controller_factory
    from .solver_api import available_solvers
    hip_cls = available_solvers.lookup(name)

    class HipAdapter(base):
        capabilities = hip_cls.capabilities
        injects_noise = True

        def __init__(self, config, grid, model):
            super().__init__(config, grid, model)
            self.ignored_keys: list = []
            cfg = Config.from_dict(config_mapping(config, self.ignored_keys))
            cfg.solver.method = name
            self._hip = controller_factory(cfg, member_overrides(model))

        @property
        def step_controller(self):
            return getattr(self._hip.solver, "step_controller", None)

        def _push(self, model, state):
            hip = self._hip
            hip.state.field = np.array(to_host(state.field), copy=True)
            hip.reservoir.set_state(tuple(np.array(to_host(r), copy=True) for r in model.reservoir.get_state()))
            hip.state.step, hip.state.t, hip.state.t0 = state.step, state.t, state.t0

        def _pull(self, model, state):
            hip = self._hip
            state.field = to_array(np.array(hip.state.field.get(), copy=True))
            model.reservoir.set_state(tuple(to_array(np.array(r.get(), copy=True))
                                            for r in hip.reservoir.get_state()))
            state.step, state.t = hip.state.step, hip.state.t

        def step(self, model, state, dt):
            """A fixed step runs as a one-tick advance, so the HIP side adds that step's Wigner
            increment (the class declares injects_noise, so the reference's loop does not);
            an adaptive trial is the HIP solver's own step."""
            self._push(model, state)
            if self.config.solver.adaptive:
                r = self._hip.solver.step(self._hip.model, self._hip.state, dt)
                report = report_cls(r.accepted, r.error, r.dt_next)
            else:
                self._hip.solver.advance(self._hip.model, self._hip.state, 1)
                report = report_cls(True, None, None)
            self._pull(model, state)
            return report

        def advance(self, model, state, n_steps):
            self._push(model, state)
            self._hip.solver.advance(self._hip.model, self._hip.state, n_steps)
            self._pull(model, state)

    HipAdapter.__name__ = HipAdapter.__qualname__ = "Hip_" + name.replace("-", "_")
    return HipAdapter
