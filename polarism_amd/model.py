"""The component bundle handed to a solver (model.py:22-49).

Unlike the reference, ``Model`` carries no right-hand side of its own: the
physics is evaluated inside the HIP kernels of the solvers.
"""
from __future__ import annotations


class Model:
    def __init__(self, grid, physics, precision, potential, pump, reservoir, boundary=None, noise=None,
                 pump_weights=None, reservoir_components=1, components=1, member_configs=None):
        self.grid = grid
        self.physics = physics
        self.precision = precision
        self.potential = potential
        self.pump = pump
        self.reservoir = reservoir
        self.boundary = boundary
        self.noise = noise
        self.terms = ()
        self.drive = None
        self.time_dependent_potential = False
        self.pump_weights = pump_weights
        self.reservoir_components = reservoir_components
        self.components = components
        self.member_configs = member_configs
