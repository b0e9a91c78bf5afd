"""Configuration dataclasses compatible with the reference's ``ps.Config``.

Section and field names, defaults and the strict unknown-key behaviour follow
polarism/config/simulation_parameters.py:10-310 so that an existing YAML file
or ``Config`` object loads unchanged; only ``solver.method`` selects one of the
HIP methods registered by this package.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass, field, fields, replace
from difflib import SequenceMatcher
from pathlib import Path
from typing import Any

import yaml


class ConfigValidationError(ValueError):
    """A configuration value is invalid (config/validation.py)."""


def suggest_keys(name: str, known, limit: int = 3) -> list[str]:
    """Known keys a mistyped ``name`` most likely meant: a case-insensitive match first, then
    the closest spellings (similarity ratio >= 0.6), best first."""
    exact = sorted(k for k in known if k.lower() == name.lower())
    scored = sorted(((SequenceMatcher(None, name, k).ratio(), k) for k in known if k not in exact),
                    key=lambda rk: (-rk[0], rk[1]))
    return (exact + [k for r, k in scored if r >= 0.6])[:limit]


def unknown_key_error(path: str, name: str, known) -> str:
    """Message for a key ``name`` that section ``path`` (empty for the top level) does not declare."""
    full = f"{path}.{name}" if path else name
    hints = suggest_keys(name, known)
    tail = f"; did you mean {', '.join(repr(h) for h in hints)}?" if hints else ""
    return f"Unknown config key '{full}'{tail}"


def config_section(name: str):
    """Declare a dataclass config section whose instances accept only their declared fields, so
    that a misspelt attribute assignment fails instead of being silently ignored."""
    def make(cls):
        cls = dataclass(cls)
        cls._section_path = name
        cls._declared = frozenset(f.name for f in fields(cls))

        def __setattr__(self, key, value):
            if key not in type(self)._declared:
                raise AttributeError(unknown_key_error(type(self)._section_path, key, type(self)._declared))
            object.__setattr__(self, key, value)

        cls.__setattr__ = __setattr__
        return cls
    return make


@config_section("grid")
class GridParameters:
    nx: int = 256
    ny: int = 256
    lx: float = 100.0
    ly: float = 100.0
    grid_type: str = "periodic"
    ndim: int = 2
    extra: dict[str, Any] = field(default_factory=dict)


@config_section("physics")
class PhysicsConstants:
    hbar: float = 0.6582119514
    m_eff: float = 0.284
    gamma_R: float = 0.005
    gamma_C: float = 0.083
    g_C: float = 0.001
    g_R: float = 0.002
    g_I: float = 0.0
    gamma_I: float = 0.001
    gamma_A: float = 0.005
    R: float = 0.02
    R_IA: float = 5e-2
    R_AI: float = 2.5e-3
    kappa: float = 0.05
    kinetic_relaxation_eta: float = 0.0
    reservoir_diffusion_I: float = 0.0
    reservoir_diffusion_A: float = 0.0
    reservoir_diffusion_R: float = 0.0
    dispersion: str = "parabolic"
    spin: str = "scalar"
    alpha_2: float = 0.0
    zeeman: float = 0.0
    te_tm: float = 0.0


@config_section("noise")
class NoiseParameters:
    enabled: bool = True
    seed: int | None = None
    initial: str = "complex_gaussian_zero_mean"
    initial_amplitude: float = 1e-3
    initial_k_cutoff_um: float | None = None
    stochastic: str = "none"
    stochastic_strength: float = 1.0


@config_section("initial_condition")
class InitialConditionParameters:
    type: str = "zero"
    amplitude: float = 1.0
    x0: float = 0.0
    y0: float = 0.0
    kx: float = 0.0
    ky: float = 0.0
    charge: int = 1
    core_width: float = 1.0
    path: str | None = None
    frame: int = -1
    extra: dict[str, Any] = field(default_factory=dict)


@config_section("boundary_condition")
class BoundaryConditionParameters:
    profile_type: str = "sin2"
    strength: float = 1.0
    absorption: str = "no-absorption"
    mask_width_percent: float = 0.2


@config_section("potential")
class PotentialParameters:
    expose_results: bool = True
    potential_type: str = "zero"
    profile: dict[str, Any] = field(default_factory=dict)
    envelope: dict[str, Any] = field(default_factory=lambda: {"type": "constant"})
    x1: float = -2.0
    y1: float = 0.0
    x2: float = 2.0
    y2: float = 0.0
    V1: float = -2.0e-3
    V2: float = -2.2e-3
    w1: float = 1.5
    w2: float = 1.5
    order: float = 2.0


@config_section("laser")
class LaserParameters:
    expose_results: bool = True
    mode: str = "single"
    config_file: str | None = None
    entry_overrides: dict[int, dict[str, Any]] = field(default_factory=dict)
    laser_type: str = "pulse-gaussian"
    profile: dict[str, Any] = field(default_factory=dict)
    envelope: dict[str, Any] = field(default_factory=lambda: {"type": "constant"})
    P0: float = 0.4
    circular_polarization: float = 0.0
    Pmax: float = 1.2
    x0: float = 0.0
    y0: float = 0.0
    sigma_space: float = 15.0
    sigma_time: float = 0.1
    pulse_separation: float = 0.6
    cutoff_sigma: float = 3.0
    delay: float = 0.0
    n_pulses: int = 0
    power_definition: str = "peak_amplitude"
    extra: dict[str, Any] = field(default_factory=dict)


@config_section("drive")
class DriveParameters:
    enabled: bool = False
    amplitude: float = 0.0
    energy: float = 0.0
    kx: float = 0.0
    ky: float = 0.0
    polarization: list[complex] = field(default_factory=lambda: [1, 0])
    profile: dict[str, Any] = field(default_factory=lambda: {"type": "uniform"})
    envelope: dict[str, Any] = field(default_factory=lambda: {"type": "constant"})


@config_section("reservoir")
class ReservoirParameters:
    expose_results: bool = True
    reservoir_type: str = "double"
    components: int | None = None
    extra: dict[str, Any] = field(default_factory=dict)


@config_section("solver")
class SolverParameters:
    total_time: float = 5.0
    dt: float = 1e-3
    method: str = "rk4-fdm-hip"
    tableau: str = "rk4"
    precision: str = "double"
    laplacian: str = "five-point"
    cuda_graphs: bool = True
    adaptive: bool = False
    rtol: float = 1e-6
    atol: float = 1e-8
    dt_max: float | None = None


@config_section("result")
class ResultParameters:
    real_time_view: bool = False
    view_member: int = 0
    real_time_refresh_interval: float = 0.1
    save_results: bool = False
    save_hdf5: bool = False
    save_json: bool = False
    save_npy: bool = False
    formats: tuple[str, ...] = ()
    observables: tuple[str, ...] = ()
    accumulators: dict[str, dict] = field(default_factory=dict)
    compression: str = "none"
    compression_level: int = 4
    save_psi: bool = False
    save_every: float | None = None
    batch_size: int = 1000
    output_directory: str = "simulation_results"
    animate: bool = False
    animation_fps: int = 8
    animation_target_seconds: int = 60
    animation_fields: tuple[str, ...] = ()
    animation_backend: str = "auto"
    animation_encoder: str = "h264_nvenc"
    animation_output: str | None = None


@config_section("compute_engine")
class ComputeEngineParameters:
    use_gpu: bool = True
    gpu_device: int = 0


@config_section("checkpoint")
class CheckpointParameters:
    every: float = 0.0
    path: str | None = None
    resume_from: str | None = None
    on_signal: bool = True


@config_section("stop")
class StopParameters:
    steady_state: bool = False
    rtol: float = 1e-6
    check_every: float = 10.0
    min_time: float = 0.0


@config_section("sweep")
class SweepParameters:
    batch_size: int | None = None
    memory_fraction: float = 0.8
    max_batch: int | None = None


@config_section("")
class Config:
    grid: GridParameters = field(default_factory=GridParameters)
    boundary_condition: BoundaryConditionParameters = field(default_factory=BoundaryConditionParameters)
    potential: PotentialParameters = field(default_factory=PotentialParameters)
    drive: DriveParameters = field(default_factory=DriveParameters)
    physics: PhysicsConstants = field(default_factory=PhysicsConstants)
    terms: dict[str, Any] = field(default_factory=dict)
    noise: NoiseParameters = field(default_factory=NoiseParameters)
    initial_condition: InitialConditionParameters = field(default_factory=InitialConditionParameters)
    laser: LaserParameters = field(default_factory=LaserParameters)
    reservoir: ReservoirParameters = field(default_factory=ReservoirParameters)
    solver: SolverParameters = field(default_factory=SolverParameters)
    result: ResultParameters = field(default_factory=ResultParameters)
    compute_engine: ComputeEngineParameters = field(default_factory=ComputeEngineParameters)
    checkpoint: CheckpointParameters = field(default_factory=CheckpointParameters)
    stop: StopParameters = field(default_factory=StopParameters)
    sweep: SweepParameters = field(default_factory=SweepParameters)

    @classmethod
    def from_dict(cls, data: dict) -> "Config":
        """Build a configuration from a nested mapping (section -> {key: value}).  Unknown
        sections or keys raise ValueError; list values of tuple-typed keys become tuples;
        ``terms`` is taken as a free mapping.  Values are not validated here."""
        defaults = cls()
        updates = {}
        for section, values in data.items():
            if section not in cls._declared:
                raise ValueError(unknown_key_error("", section, cls._declared))
            if not isinstance(values, dict):
                raise ValueError(f"Config section '{section}' must be a mapping")
            if section == "terms":
                updates["terms"] = dict(values)
                continue
            current = getattr(defaults, section)
            declared = type(current)._declared
            stray = [k for k in values if k not in declared]
            if stray:
                raise ValueError(unknown_key_error(section, stray[0], declared))
            coerced = {k: tuple(v) if isinstance(getattr(current, k), tuple) and isinstance(v, list) else v
                       for k, v in values.items()}
            updates[section] = replace(current, **coerced)
        return replace(defaults, **updates)

    @classmethod
    def from_yaml(cls, path) -> "Config":
        """Load nested YAML (simulation_parameters.py:275-302) and validate it."""
        from .validation import validate_config
        with open(path, encoding="utf-8") as handle:
            data = yaml.safe_load(handle)
        if not isinstance(data, dict):
            raise ValueError("Config YAML must be a mapping")
        cfg = cls.from_dict(data)
        validate_config(cfg)
        return cfg

    def to_yaml(self, path) -> Path:
        out = Path(path)
        with out.open("w", encoding="utf-8") as handle:
            yaml.safe_dump(asdict(self), handle, sort_keys=False)
        return out


FLOAT_PHYSICS_FIELDS = frozenset(f.name for f in fields(PhysicsConstants) if f.type in (float, "float"))
