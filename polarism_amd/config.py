"""Configuration dataclasses compatible with the reference's ``ps.Config``.

Section and field names, defaults and the strict unknown-key behaviour follow
polarism/config/simulation_parameters.py:10-310 so that an existing YAML file
or ``Config`` object loads unchanged; only ``solver.method`` selects one of the
HIP methods registered by this package.
"""
from __future__ import annotations

from dataclasses import asdict, dataclass, field, fields
from difflib import get_close_matches
from pathlib import Path
from typing import Any

import yaml


class ConfigValidationError(ValueError):
    """A configuration value is invalid (config/validation.py)."""


def unknown_field_message(owner: str, name: str, known) -> str:
    known = list(known)
    matches = [k for k in known if k.lower() == name.lower()]
    matches += [k for k in get_close_matches(name, known, n=3, cutoff=0.6) if k not in matches]
    hint = f"; did you mean {', '.join(repr(m) for m in matches)}?" if matches else ""
    key = name if owner == "Config" else f"{owner}.{name}"
    return f"Unknown config key '{key}'{hint}"


class StrictFields:
    """Reject attributes that are not declared fields (simulation_parameters.py:19-27)."""

    def __setattr__(self, name, value):
        if name not in self.__dataclass_fields__:
            owner = _SECTION_NAMES.get(type(self), type(self).__name__)
            raise AttributeError(unknown_field_message(owner, name, self.__dataclass_fields__))
        object.__setattr__(self, name, value)


@dataclass
class GridParameters(StrictFields):
    nx: int = 256
    ny: int = 256
    lx: float = 100.0
    ly: float = 100.0
    grid_type: str = "periodic"
    ndim: int = 2
    extra: dict[str, Any] = field(default_factory=dict)


@dataclass
class PhysicsConstants(StrictFields):
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


@dataclass
class NoiseParameters(StrictFields):
    enabled: bool = True
    seed: int | None = None
    initial: str = "complex_gaussian_zero_mean"
    initial_amplitude: float = 1e-3
    initial_k_cutoff_um: float | None = None
    stochastic: str = "none"
    stochastic_strength: float = 1.0


@dataclass
class InitialConditionParameters(StrictFields):
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


@dataclass
class BoundaryConditionParameters(StrictFields):
    profile_type: str = "sin2"
    strength: float = 1.0
    absorption: str = "no-absorption"
    mask_width_percent: float = 0.2


@dataclass
class PotentialParameters(StrictFields):
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


@dataclass
class LaserParameters(StrictFields):
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


@dataclass
class DriveParameters(StrictFields):
    enabled: bool = False
    amplitude: float = 0.0
    energy: float = 0.0
    kx: float = 0.0
    ky: float = 0.0
    polarization: list[complex] = field(default_factory=lambda: [1, 0])
    profile: dict[str, Any] = field(default_factory=lambda: {"type": "uniform"})
    envelope: dict[str, Any] = field(default_factory=lambda: {"type": "constant"})


@dataclass
class ReservoirParameters(StrictFields):
    expose_results: bool = True
    reservoir_type: str = "double"
    components: int | None = None
    extra: dict[str, Any] = field(default_factory=dict)


@dataclass
class SolverParameters(StrictFields):
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


@dataclass
class ResultParameters(StrictFields):
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


@dataclass
class ComputeEngineParameters(StrictFields):
    use_gpu: bool = True
    gpu_device: int = 0


@dataclass
class CheckpointParameters(StrictFields):
    every: float = 0.0
    path: str | None = None
    resume_from: str | None = None
    on_signal: bool = True


@dataclass
class StopParameters(StrictFields):
    steady_state: bool = False
    rtol: float = 1e-6
    check_every: float = 10.0
    min_time: float = 0.0


@dataclass
class SweepParameters(StrictFields):
    batch_size: int | None = None
    memory_fraction: float = 0.8
    max_batch: int | None = None


@dataclass
class Config(StrictFields):
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
        cfg = cls()
        sections = {f.name for f in fields(cls)}
        for section, values in data.items():
            if section not in sections:
                raise ValueError(unknown_field_message("Config", section, sections))
            if not isinstance(values, dict):
                raise ValueError(f"Config section '{section}' must be a mapping")
            if section == "terms":
                cfg.terms = values
                continue
            target = getattr(cfg, section)
            known = {f.name for f in fields(target)}
            for name, value in values.items():
                if name not in known:
                    raise ValueError(unknown_field_message(section, name, known))
                if isinstance(getattr(target, name), tuple) and isinstance(value, list):
                    value = tuple(value)
                setattr(target, name, value)
        return cfg

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


_SECTION_NAMES = {f.default_factory: f.name for f in fields(Config)
                  if isinstance(f.default_factory, type) and issubclass(f.default_factory, StrictFields)}

FLOAT_PHYSICS_FIELDS = frozenset(f.name for f in fields(PhysicsConstants) if f.type in (float, "float"))
