"""Configuration validation for the HIP methods (config/validation.py:54-71, 195-242)."""
from __future__ import annotations

import math

from .config import ConfigValidationError

_VALID_PRECISIONS = ("single", "double")
_TABLEAUS = {"rk4", "dp5", "dop853"}


def validate_config(cfg) -> None:
    s = cfg.solver
    if cfg.grid.ndim != 2:
        raise ConfigValidationError("grid.ndim must be 2 for the HIP solvers")
    if cfg.grid.nx < 1 or cfg.grid.ny < 1:
        raise ConfigValidationError("grid.nx and grid.ny must be positive")
    if cfg.grid.lx <= 0 or cfg.grid.ly <= 0:
        raise ConfigValidationError("grid.lx and grid.ly must be positive")
    if s.tableau not in _TABLEAUS:
        raise ConfigValidationError("solver.tableau must be one of rk4, dp5, dop853")
    if s.dt <= 0.0:
        raise ConfigValidationError(f"solver.dt must be > 0, got {s.dt}")
    if s.total_time <= 0.0:
        raise ConfigValidationError(f"solver.total_time must be > 0, got {s.total_time}")
    if s.precision not in _VALID_PRECISIONS:
        raise ConfigValidationError(f"solver.precision must be one of {_VALID_PRECISIONS}, got '{s.precision}'")
    if type(s.adaptive) is not bool:
        raise ConfigValidationError("solver.adaptive must be a boolean")
    if cfg.physics.spin not in ("scalar", "spinor"):
        raise ConfigValidationError("physics.spin must be 'scalar' or 'spinor'")
    if cfg.noise.stochastic not in ("none", "wigner"):
        raise ConfigValidationError("noise.stochastic must be 'none' or 'wigner'")
    if cfg.stop.steady_state and cfg.stop.check_every <= 0:
        raise ConfigValidationError("stop.check_every must be > 0")
    _validate_accumulators(cfg.result.accumulators)
    if s.adaptive:
        from .solver_api import solver_capabilities
        from .tableaus import TABLEAUS
        if not solver_capabilities(s.method).adaptive:
            raise ConfigValidationError(f"solver.method={s.method!r} does not support solver.adaptive")
        if TABLEAUS[s.tableau].embedded_order is None:
            raise ConfigValidationError("solver.adaptive requires an embedded solver.tableau")
        if cfg.noise.enabled and cfg.noise.stochastic != "none":
            raise ConfigValidationError("solver.adaptive requires stochastic noise off")
        for name, positive in (("rtol", True), ("atol", False), ("dt_max", True)):
            value = getattr(s, name)
            if value is None and name == "dt_max":
                continue
            if (isinstance(value, bool) or not isinstance(value, (int, float)) or not math.isfinite(value)
                    or (value <= 0 if positive else value < 0)):
                raise ConfigValidationError(f"solver.{name} must be {'> 0' if positive else '>= 0'} and finite")
        if s.precision == "single" and s.rtol < 1e-5:
            raise ConfigValidationError("solver.rtol must be >= 1e-5 in single precision")


def _validate_accumulators(accumulators) -> None:
    """result.accumulators maps registered names to parameter mappings (config/validation.py:274-287)."""
    from .accumulators import available_accumulators
    if not isinstance(accumulators, dict):
        raise ConfigValidationError("result.accumulators must map names to parameter mappings")
    for name, params in accumulators.items():
        try:
            accumulator = available_accumulators.lookup(name)
        except ValueError as exc:
            raise ConfigValidationError(f"result.accumulators: {exc}") from exc
        if not isinstance(params, dict):
            raise ConfigValidationError(f"result.accumulators.{name} must be a parameter mapping")
        unknown = set(params) - {"start", "stop", "every"} - accumulator.parameter_keys
        if unknown:
            raise ConfigValidationError(
                f"result.accumulators.{name}: unknown keys {sorted(unknown)}; "
                f"allowed: {sorted({'start', 'stop', 'every'} | accumulator.parameter_keys)}")
