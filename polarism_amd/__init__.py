"""MI355X-native driven-dissipative Gross-Pitaevskii time stepper.

Registers HIP methods behind a solver registry compatible with Polarism's
(``rk4-fdm-hip`` and friends); select one with ``config.solver.method``.
"""
from .config import Config, ConfigValidationError
from .solver_api import (ACCEPTED, AbstractSolver, SolverCapabilities, SolverCompatibilityError, StepReport,
                         available_solvers, check_solver_compatibility, create_solver, register_solver)
from . import solvers  # noqa: F401  (registration)
from .controller import SimulationController, apply_overrides

__version__ = "0.1.0"

__all__ = ["Config", "ConfigValidationError", "ACCEPTED", "AbstractSolver", "SolverCapabilities",
           "SolverCompatibilityError", "StepReport", "available_solvers", "check_solver_compatibility",
           "create_solver", "register_solver", "SimulationController", "apply_overrides"]
