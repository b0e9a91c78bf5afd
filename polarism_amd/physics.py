"""Per-member physics values (config/batch_physics.py:17-64).

``BatchPhysics.get(name)`` returns the shared scalar, or a float64 host
column of one value per member when the sweep varies that constant.
The HIP solvers read these columns into the device member table.
"""
from __future__ import annotations

import numpy as np

from .config import FLOAT_PHYSICS_FIELDS


def _member_column(name: str, value) -> np.ndarray:
    """One float64 value per member for physics constant ``name``, or ValueError."""
    if name not in FLOAT_PHYSICS_FIELDS:
        raise ValueError(f"Unknown or non-float batch_physics field: {name!r}")
    col = np.asarray(value, dtype=np.float64)
    if col.ndim != 1 or col.size == 0:
        raise ValueError(f"batch_physics.{name} must be a nonempty 1-D array")
    if not np.isfinite(col).all():
        raise ValueError(f"batch_physics.{name} must contain finite values")
    return col


class BatchPhysics:
    def __init__(self, physics, values=None, precision="double"):
        self.physics = physics
        self.precision = precision
        self.host: dict[str, np.ndarray] = {k: _member_column(k, v) for k, v in (values or {}).items()}
        lengths = {k: c.size for k, c in self.host.items()}
        if len(set(lengths.values())) > 1:
            first = next(iter(lengths.values()))
            odd = next(k for k, n in lengths.items() if n != first)
            raise ValueError(f"batch_physics.{odd} has length {lengths[odd]}; expected {first}")
        self.size = next(iter(lengths.values()), None)

    def get(self, name):
        """Return the per-member column if the value varies, else the scalar."""
        if name in self.host:
            return self.host[name]
        return getattr(self.physics, name, 0.0)

    def member(self, name, m: int) -> float:
        """The value of ``name`` for member ``m``."""
        if name in self.host:
            return float(self.host[name][m])
        return float(getattr(self.physics, name, 0.0))

    def any_nonzero(self, name) -> bool:
        if name in self.host:
            return bool(np.any(self.host[name]))
        return getattr(self.physics, name, 0.0) != 0.0

    def check(self, shape, components=1):
        if len(shape) != 4 or shape[1] != components:
            raise ValueError(f"batch_physics requires canonical field shape; got {shape}")
        if self.size is not None and shape[0] != self.size:
            raise ValueError(f"batch_physics length {self.size} requires state batch size {self.size}; got {shape[0]}")
