"""Device-resident reservoir state holders.

Each holder keeps its fields as ``(B, C_res, ny, nx)`` DeviceArrays and
exposes the reference interface the solvers and users rely on
(reservoir/single_reservoir.py:22-111, double_reservoir.py:171-292,
quadratic_double_reservoir.py:41-226, none_reservoir.py):
``state_names``, ``get_state``, ``set_state`` (clamps at zero),
``get_active_density``, ``get_inactive_density``.  The derivatives
themselves are evaluated inside the HIP stage kernels.
"""
from __future__ import annotations

import numpy as np

from .device import DeviceArray


def _real(precision):
    return np.float32 if precision == "single" else np.float64


class DeviceReservoir:
    kind = "none"
    state_names: tuple[str, ...] = ()

    def __init__(self, grid, precision="double", batch_size=1, components=1):
        self.precision = precision
        self.components = components
        self.shape = (batch_size, components, *grid.shape)
        self._fields = tuple(DeviceArray.zeros(self.shape, _real(precision)) for _ in self.state_names)

    # -- reference interface ------------------------------------------------
    def get_state(self):
        return self._fields

    def set_state(self, state) -> None:
        """Replace the fields; host arrays are clamped at zero (single_reservoir.py:70-73)."""
        if len(state) != len(self.state_names):
            raise ValueError(f"{self.kind} reservoir state has {len(self.state_names)} fields")
        fields = []
        for value in state:
            if isinstance(value, DeviceArray):
                if value.shape != self.shape or value.dtype != np.dtype(_real(self.precision)):
                    raise ValueError(f"reservoir field must be {self.shape} {_real(self.precision).__name__}")
                fields.append(value)
            else:
                host = np.maximum(np.asarray(value, dtype=_real(self.precision)), 0)
                host = np.broadcast_to(host.reshape(host.shape if host.ndim == 4 else (1, 1, *host.shape[-2:])),
                                       self.shape)
                fields.append(DeviceArray.from_numpy(host))
        self._fields = tuple(fields)

    def assign_device(self, fields) -> None:
        """Adopt solver-owned device fields (already clamped by the kernel)."""
        self._fields = tuple(fields)

    def get_active_density(self, state=None):
        state = self._fields if state is None else state
        return state[0] if state else None

    def get_inactive_density(self, state=None):
        state = self._fields if state is None else state
        return state[1] if len(state) > 1 else None

    def get_reservoir_density(self):
        return self.get_active_density()


class NoReservoir(DeviceReservoir):
    kind = "none"
    state_names = ()


class SingleReservoir(DeviceReservoir):
    kind = "single"
    state_names = ("nR",)


class DoubleReservoir(DeviceReservoir):
    kind = "double"
    state_names = ("nA", "nI")


class QuadraticDoubleReservoir(DeviceReservoir):
    kind = "quadratic-double"
    state_names = ("nR", "nI")


RESERVOIRS = {"none": NoReservoir, "single": SingleReservoir, "double": DoubleReservoir,
              "quadratic-double": QuadraticDoubleReservoir}


def create_reservoir(cfg, physics, grid, precision, batch_size=1):
    kind = cfg.reservoir_type
    if kind not in RESERVOIRS:
        raise ValueError(f"Unknown reservoir '{kind}'. Available: {sorted(RESERVOIRS)}")
    if physics.reservoir_diffusion_R or physics.reservoir_diffusion_A or physics.reservoir_diffusion_I:
        raise ValueError("reservoir diffusion is not supported by the HIP solvers")
    components = cfg.components or (2 if physics.spin == "spinor" else 1)
    return RESERVOIRS[kind](grid, precision, batch_size, components)


def reservoir_rates(kind: str, get) -> tuple[float, float, float, float]:
    """Pack the reservoir constants of one member into (res_a, res_b, res_c, res_d).

    ``get(name)`` returns the member's float value.  Sums of rates are taken
    in double before they reach the device, as the reference adds Python
    floats before touching the array (double_reservoir.py:268-269).
    """
    if kind == "single":
        return (get("gamma_R"), 0.0, 0.0, 0.0)
    if kind == "double":
        return (get("gamma_I") + get("R_IA"), get("R_AI"), get("R_IA"), get("gamma_A") + get("R_AI"))
    if kind == "quadratic-double":
        return (get("kappa"), get("gamma_R"), get("gamma_I"), 0.0)
    return (0.0, 0.0, 0.0, 0.0)
