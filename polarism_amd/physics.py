"""Per-member physics values (config/batch_physics.py:17-64).

``BatchPhysics.get(name)`` returns the shared scalar, or a float64 host
column of one value per member when the sweep varies that constant.
The HIP solvers read these columns into the device member table.
"""
from __future__ import annotations

import numpy as np

from .config import FLOAT_PHYSICS_FIELDS


class BatchPhysics:
    def __init__(self, physics, values=None, precision="double"):
        self.physics = physics
        self.host: dict[str, np.ndarray] = {}
        self.size = None
        for name, value in (values or {}).items():
            if name not in FLOAT_PHYSICS_FIELDS:
                raise ValueError(f"Unknown or non-float batch_physics field: {name!r}")
            arr = np.asarray(value, dtype=np.float64)
            if arr.ndim != 1 or arr.size == 0:
                raise ValueError(f"batch_physics.{name} must be a nonempty 1-D array")
            if self.size is not None and arr.size != self.size:
                raise ValueError(f"batch_physics.{name} has length {arr.size}; expected {self.size}")
            if not np.all(np.isfinite(arr)):
                raise ValueError(f"batch_physics.{name} must contain finite values")
            self.size = arr.size
            self.host[name] = arr
        self.precision = precision

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
