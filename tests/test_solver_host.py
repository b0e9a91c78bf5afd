"""Host logic of the drop-in boundary: registry, capabilities and configuration checks (no GPU)."""
import warnings

import pytest

import polarism_amd  # noqa: F401  (registers the HIP methods)
from polarism_amd.config import ConfigValidationError
from polarism_amd.solver_api import (ACCEPTED, Registry, SolverCompatibilityError, StepReport, available_solvers,
                                     check_solver_compatibility, register_solver, solver_capabilities)
from polarism_amd.validation import validate_config
from tests.helpers import quickstart_config

HIP_METHODS = ("rk4-fdm-hip", "split-step-fft-hip", "split-step-fft-kouter-hip", "ifrk-fft-hip", "ifrk4-fft-hip",
               "ip-rk4-hip")


def test_every_hip_method_is_registered_with_declared_capabilities():
    for name in HIP_METHODS:
        assert name in available_solvers
        with warnings.catch_warnings():
            warnings.simplefilter("error")
            caps = solver_capabilities(name)
        assert caps.description


def test_registry_warns_and_overrides_on_duplicate_names():
    """registry.py:18-30: re-registering a name warns and the newer class wins."""
    reg = Registry("solver")

    class A: pass

    class B: pass

    reg.register("x")(A)
    with pytest.warns(UserWarning, match="already registered"):
        reg.register("x")(B)
    assert reg.lookup("x") is B
    with pytest.raises(ValueError, match="Unknown solver 'y'"):
        reg.lookup("y")


def test_plugin_without_capabilities_gets_conservative_defaults():
    @register_solver("host-test-plugin")
    class Plugin:
        pass
    try:
        with pytest.warns(UserWarning, match="does not declare capabilities"):
            caps = solver_capabilities("host-test-plugin")
        assert caps.components == frozenset({1}) and not caps.adaptive
    finally:
        del available_solvers["host-test-plugin"]


def test_fixed_step_report_is_the_shared_accepted():
    assert ACCEPTED.accepted and ACCEPTED.error is None
    assert StepReport(accepted=False, error=2.0).accepted is False


@pytest.mark.parametrize("method,ok", [("rk4-fdm-hip", True), ("split-step-fft-hip", False),
                                       ("ifrk-fft-hip", False), ("ip-rk4-hip", False)])
def test_closed_interval_grids_only_on_the_finite_difference_method(method, ok):
    cfg = quickstart_config(solver__method=method, grid__grid_type="closed-interval")
    if ok:
        check_solver_compatibility(cfg)
    else:
        with pytest.raises(SolverCompatibilityError, match="closed-interval"):
            check_solver_compatibility(cfg)


@pytest.mark.parametrize("method", ["ifrk4-fft-hip", "ip-rk4-hip", "rk4-fdm-hip", "split-step-fft-hip"])
def test_adaptive_stepping_only_on_the_lawson_family(method):
    cfg = quickstart_config(solver__method=method, solver__adaptive=True, solver__tableau="dp5")
    with pytest.raises(ConfigValidationError, match="does not support solver.adaptive"):
        validate_config(cfg)
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__adaptive=True, solver__tableau="dp5")
    validate_config(cfg)


def test_single_precision_adaptive_rtol_floor():
    """config/validation.py:241-242: rtol below 1e-5 is rejected in single precision only."""
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__adaptive=True, solver__tableau="dp5",
                            solver__precision="single", solver__rtol=1e-6)
    with pytest.raises(ConfigValidationError, match="1e-5"):
        validate_config(cfg)
    cfg.solver.precision = "double"
    validate_config(cfg)


def test_adaptive_needs_an_embedded_tableau():
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__adaptive=True, solver__tableau="rk4")
    with pytest.raises(ConfigValidationError, match="embedded"):
        validate_config(cfg)


@pytest.mark.parametrize("method", HIP_METHODS)
def test_every_hip_method_takes_spinor_fields(method):
    """Like the reference methods they replace, every HIP method accepts psi+ / psi- fields."""
    cfg = quickstart_config(solver__method=method, physics__spin="spinor")
    check_solver_compatibility(cfg)


def test_kinetic_relaxation_needs_a_method_that_evaluates_it():
    for method in HIP_METHODS:
        cfg = quickstart_config(solver__method=method, physics__kinetic_relaxation_eta=0.02)
        if solver_capabilities(method).kinetic_relaxation:
            check_solver_compatibility(cfg)
        else:
            with pytest.raises(SolverCompatibilityError, match="kinetic_relaxation_eta"):
                check_solver_compatibility(cfg)
