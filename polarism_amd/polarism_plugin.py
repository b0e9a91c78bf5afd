"""Polarism plugin module: registers the HIP methods in the reference's solver registry.

Copy (or link) this file into ``./polarism_plugins/`` of a Polarism run directory; Polarism
imports it at start-up (polarism/plugins.py:12-41).  It needs ``polarism_amd`` importable and
``libpgpe_hip.so`` built.  The wrapping itself lives in ``polarism_amd.plugin_adapter``.
"""
POLARISM_PLUGIN_API = 2

from polarism.compute_engine import compute_engine  # noqa: E402
from polarism.solver.abstract_solver import AbstractSolver, StepReport  # noqa: E402
from polarism.solver.solver_registry import register_solver  # noqa: E402

from polarism_amd import solvers as _hip_solvers  # noqa: E402,F401  (registers the HIP methods)
from polarism_amd.plugin_adapter import HIP_METHODS, make_hip_solver  # noqa: E402

for _name in HIP_METHODS:
    register_solver(_name)(make_hip_solver(_name, base=AbstractSolver, report_cls=StepReport,
                                           to_host=compute_engine.to_cpu,
                                           to_array=lambda a: compute_engine.xp.asarray(a)))
