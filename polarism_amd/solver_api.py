"""The solver plugin contract: registry, capabilities, step reports, base class.

Mirrors the reference's drop-in boundary so that a registered method is
selected by ``config.solver.method`` exactly as in Polarism:

* ``Registry`` / ``register_solver`` / ``create_solver`` --
  polarism/registry.py:11-37, solver/solver_registry.py:12-17,
  solver/create_solver.py:7-9
* ``StepReport``, ``ACCEPTED``, ``SolverCapabilities``, ``AbstractSolver`` --
  solver/abstract_solver.py:11-101
* ``check_solver_compatibility`` / ``SolverCompatibilityError`` --
  solver/solver_compatibility.py:13-137
"""
from __future__ import annotations

import warnings
from abc import ABC, abstractmethod
from collections.abc import Callable
from dataclasses import dataclass
from typing import ClassVar, Generic, TypeVar

T = TypeVar("T")


class Registry(dict, Generic[T]):
    """Map component names to implementations; re-registration warns and overrides."""

    def __init__(self, kind: str) -> None:
        super().__init__()
        self.kind = kind

    def register(self, name: str) -> Callable[[T], T]:
        def decorator(component: T) -> T:
            existing = self.get(name)
            if existing is not None and existing is not component:
                warnings.warn(f"{self.kind.title()} '{name}' is already registered; overriding {existing!r}.",
                              stacklevel=2)
            self[name] = component
            return component
        return decorator

    def lookup(self, name: str) -> T:
        try:
            return self[name]
        except KeyError as exc:
            raise ValueError(f"Unknown {self.kind} '{name}'. Available: {sorted(self)}") from exc


@dataclass(frozen=True)
class StepReport:
    """Outcome of a solver step (abstract_solver.py:11-17)."""
    accepted: bool = True
    error: float | None = None
    dt_next: float | None = None


ACCEPTED = StepReport()

_BUILTIN_RESERVOIRS = frozenset({"single", "double", "quadratic-double", "none"})


@dataclass(frozen=True)
class SolverCapabilities:
    """Features a solver accepts (abstract_solver.py:26-57)."""
    components: frozenset = frozenset({1})
    ndim: frozenset = frozenset({2})
    grid_types: frozenset = frozenset({"periodic"})
    reservoirs: frozenset = _BUILTIN_RESERVOIRS
    custom_reservoirs: str = "quantitative"
    potential: bool = True
    kinetic_relaxation: bool = False
    dispersion: frozenset = frozenset({"parabolic"})
    terms: bool = True
    noise: frozenset = frozenset({"none", "wigner"})
    adaptive: bool = False
    coherent_drive: bool = False
    time_dependent_potential: bool = False
    quantitative_reservoirs: frozenset = _BUILTIN_RESERVOIRS
    description: str = ""

    def __post_init__(self):
        for name in ("components", "ndim", "grid_types", "reservoirs", "dispersion", "noise",
                     "quantitative_reservoirs"):
            object.__setattr__(self, name, frozenset(getattr(self, name)))


DEFAULT_CAPABILITIES = SolverCapabilities(description="Plugin solver with undeclared capabilities.")

available_solvers: Registry = Registry("solver")


def register_solver(name: str):
    """Return a decorator that registers a solver class under ``name``."""
    return available_solvers.register(name)


def create_solver(config, grid, model):
    """Instantiate the registered implementation of ``config.solver.method``."""
    return available_solvers.lookup(config.solver.method)(config, grid, model)


class SolverCompatibilityError(RuntimeError):
    """The solver cannot run this configuration."""


def solver_capabilities(name: str) -> SolverCapabilities:
    component = available_solvers.lookup(name)
    caps = getattr(component, "capabilities", DEFAULT_CAPABILITIES)
    if caps is DEFAULT_CAPABILITIES:
        warnings.warn(f"Solver '{name}' does not declare capabilities; using conservative defaults.",
                      UserWarning, stacklevel=2)
    return caps


def check_solver_compatibility(cfg) -> None:
    """Raise SolverCompatibilityError if the method cannot run ``cfg``."""
    solver = cfg.solver.method
    caps = solver_capabilities(solver)
    if cfg.grid.ndim not in caps.ndim:
        raise SolverCompatibilityError(f"solver.method={solver!r} does not support grid.ndim={cfg.grid.ndim}")
    components = 2 if cfg.physics.spin == "spinor" else 1
    if components not in caps.components:
        raise SolverCompatibilityError(f"Solver '{solver}' does not support {components} components")
    if cfg.grid.grid_type not in caps.grid_types:
        raise SolverCompatibilityError(
            f"Solver '{solver}' does not support grid_type='{cfg.grid.grid_type}'. "
            f"Supported: {sorted(caps.grid_types)}. Hint: {caps.description}")
    rtype = cfg.reservoir.reservoir_type
    if rtype not in caps.reservoirs:
        raise SolverCompatibilityError(
            f"Solver '{solver}' does not support reservoir_type='{rtype}'. "
            f"Supported: {sorted(caps.reservoirs)}. Hint: {caps.description}")
    if "*" not in caps.dispersion and cfg.physics.dispersion not in caps.dispersion:
        raise SolverCompatibilityError(
            f"Solver '{solver}' does not support physics.dispersion='{cfg.physics.dispersion}'.")
    if cfg.terms and not caps.terms:
        raise SolverCompatibilityError(f"Solver '{solver}' does not support extra terms {sorted(cfg.terms)}.")
    if cfg.drive.enabled and not caps.coherent_drive:
        raise SolverCompatibilityError(f"Solver '{solver}' does not support drive.enabled=True.")
    if cfg.potential.envelope.get("type") != "constant" and not caps.time_dependent_potential:
        raise SolverCompatibilityError(f"Solver '{solver}' does not support time-dependent potential.envelope.")
    if cfg.noise.enabled and cfg.noise.stochastic not in caps.noise:
        raise SolverCompatibilityError(
            f"Solver '{solver}' does not support noise.stochastic='{cfg.noise.stochastic}'")
    if cfg.physics.kinetic_relaxation_eta != 0.0 and not caps.kinetic_relaxation:
        raise SolverCompatibilityError(
            f"Solver '{solver}' does not evaluate kinetic_relaxation_eta; set it to 0 or switch solver.")
    if cfg.potential.potential_type != "zero" and not caps.potential:
        raise SolverCompatibilityError(f"Solver '{solver}' does not support a nonzero potential.")


class AbstractSolver(ABC):
    """Base class of every solver (abstract_solver.py:60-101).

    ``step`` advances one fixed step (``state.step += 1``,
    ``state.t = state.t0 + state.step*dt``) or performs one adaptive trial that
    only commits when error <= 1.  ``advance`` runs ``n_steps`` ticks and
    applies ``model.noise`` after each accepted step unless the solver sets
    ``injects_noise``.
    """

    capabilities: ClassVar[SolverCapabilities] = DEFAULT_CAPABILITIES
    injects_noise: ClassVar[bool] = False

    def __init__(self, config, grid, model):
        self.config = config
        self.physics = config.physics
        self.batch_physics = model.physics

    @abstractmethod
    def step(self, model, state, dt) -> StepReport:
        raise NotImplementedError

    def adaptive_control(self):
        """The solver's PI controller (``step_controller``), created on first use."""
        from .adaptive import PIControl
        if getattr(self, "step_controller", None) is None:
            self.step_controller = PIControl.for_solver(self)
        return self.step_controller

    def run_trials(self, model, state, target: float) -> None:
        """Adaptive trials until ``state.t`` reaches ``target``; solvers with device-side
        control override this."""
        from .adaptive import host_trials
        host_trials(self.adaptive_control(), self, model, state, target)

    def advance(self, model, state, n_steps: int) -> None:
        if self.config.solver.adaptive:
            target = state.t0 + (state.step + n_steps) * self.config.solver.dt
            self.run_trials(model, state, target)
            state.step += n_steps
            state.t = state.t0 + state.step * self.config.solver.dt
            return
        for _ in range(n_steps):
            completed = state.step
            report = self.step(model, state, self.config.solver.dt)
            if report.accepted and model.noise is not None and not self.injects_noise:
                model.noise.apply_to_state(state, model.reservoir, completed)
