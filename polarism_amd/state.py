"""Canonical simulation state on the device (simulation_state.py:84-260).

``field`` is the ``(B, C, ny, nx)`` complex DeviceArray; ``psi`` is the host
NumPy view of the public shape (``(ny, nx)`` for one scalar member,
``(B, ny, nx)`` batched, with a component axis for spinors).  ``field_k`` is
the spectral cache a spectral solver may leave instead of ``field``.
"""
from __future__ import annotations

import numpy as np

from .device import DeviceArray
from .noise import make_initial_psi


def complex_dtype(precision):
    return np.complex64 if precision == "single" else np.complex128


class SimulationState:
    def __init__(self, grid, eps=1e-3, precision="double", *, mode="complex_gaussian_zero_mean",
                 seed=None, batch_size=None, components=1, initial_fields=None):
        if batch_size is not None and (isinstance(batch_size, bool) or not isinstance(batch_size, int)
                                       or batch_size < 1):
            raise ValueError("batch_size must be a positive integer")
        if components not in (1, 2):
            raise ValueError("components must be 1 or 2")
        self.batch_size = 1 if batch_size is None else batch_size
        self.batched = batch_size is not None
        self.components = components
        self.precision = precision
        self._spatial_shape = grid.shape
        if initial_fields is None:
            psi = make_initial_psi(grid.shape, eps, mode, seed, precision)
            if components == 2:
                other = make_initial_psi(grid.shape, eps, mode, None if seed is None else seed + 1, precision)
                psi = np.stack((psi, other))
            host = np.broadcast_to(psi.reshape(1, components, *grid.shape), self._canonical_shape)
        else:
            host = np.asarray(initial_fields, dtype=complex_dtype(precision)).reshape(self._canonical_shape)
        self._field = DeviceArray.from_numpy(host, dtype=complex_dtype(precision))
        self._field_k = None
        self.t = 0.0
        self.t0 = 0.0
        self.step = 0

    @property
    def _canonical_shape(self):
        return (self.batch_size, self.components, *self._spatial_shape)

    @property
    def _public_shape(self):
        if self.components == 2:
            return (self.batch_size, 2, *self._spatial_shape) if self.batched else (2, *self._spatial_shape)
        return (self.batch_size, *self._spatial_shape) if self.batched else self._spatial_shape

    @property
    def field(self) -> DeviceArray:
        if self._field is None:
            raise ValueError("SimulationState.field is unavailable (only the spectral cache is set)")
        return self._field

    @field.setter
    def field(self, value) -> None:
        if value is None:
            self._field = None
            return
        if tuple(value.shape) != self._canonical_shape:
            raise ValueError(f"field shape must be {self._canonical_shape}, got {tuple(value.shape)}")
        if not isinstance(value, DeviceArray):
            value = DeviceArray.from_numpy(np.asarray(value, dtype=complex_dtype(self.precision)))
        self._field = value
        self._field_k = None

    @property
    def field_k(self):
        return self._field_k

    @field_k.setter
    def field_k(self, value) -> None:
        if value is not None and tuple(value.shape) != self._canonical_shape:
            raise ValueError(f"field_k shape must be {self._canonical_shape}")
        self._field_k = value

    def invalidate_psi_k(self) -> None:
        self._field_k = None

    @property
    def psi(self) -> np.ndarray:
        """Host copy of the field in the public shape."""
        return self.field.get().reshape(self._public_shape)

    @psi.setter
    def psi(self, value) -> None:
        value = np.asarray(value)
        if value.shape != self._public_shape:
            raise ValueError(f"psi shape must be {self._public_shape}, got {value.shape}")
        self.field = value.reshape(self._canonical_shape)

    def get_simulation_state(self):
        return self.psi, self.t
