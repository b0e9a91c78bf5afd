"""``rk4-fdm-hip``: the stage-coupled RK4 finite-difference stepper on gfx950.

Drop-in for the reference's ``rk4-fdm`` / ``rk4-cuda`` methods
(solver/rk4_fdm_solver.py:14-77, solver/rk4_cuda_solver.py:2197-3642):
the same RK4 stages, reservoir clamp and noise semantics, with one
``pgpe_rk4_step`` launch per step over every member of the batch (all four
stages on halo tiles) for scalar fields without a potential, and four fused
``pgpe_rk4_stage`` launches per step otherwise (same result).  Chunks
of up to 1000 steps are captured as HIP graphs whose pump coefficients and
noise step base live in device buffers, so replaying a chunk needs no host
synchronisation inside the chunk.
"""
from __future__ import annotations

import numpy as np

from ..autotune import candidates, tune_rk4
from ..device import DeviceArray
from ..engine import STRIP_MAX_POINTS, Rk4Engine, spinor_whole_step_applies
from ..solver_api import ACCEPTED, SolverCapabilities, SolverCompatibilityError, register_solver
from .base import HipSolverBase

_ALL_RES = frozenset({"single", "double", "quadratic-double", "none"})


@register_solver("rk4-fdm-hip")
class RK4FDMHipSolver(HipSolverBase):
    capabilities = SolverCapabilities(
        components=frozenset({1, 2}), ndim=frozenset({2}), grid_types=frozenset({"periodic", "closed-interval"}),
        reservoirs=_ALL_RES, custom_reservoirs="unsupported", potential=True,
        kinetic_relaxation=True, dispersion=frozenset({"parabolic"}), terms=False,
        noise=frozenset({"none", "wigner"}), quantitative_reservoirs=_ALL_RES,
        description="Stage-coupled FDM RK4 on gfx950: one fused HIP launch per RK stage over all "
                    "members, LDS halo tiles, in-kernel Wigner noise, HIP-graph chunks; scalar or spinor "
                    "(TE-TM 9-point, Zeeman, alpha_2). 2-D periodic or closed-interval grids, 5-point or "
                    "isotropic 9-point Laplacian, static complex potential.")
    injects_noise = True   # advance() adds the Wigner increment inside the final stage kernel
    max_graph_chunk = 1000

    def __init__(self, config, grid, model):
        super().__init__(config, grid, model)
        self.engine = None
        self._const_coef = None
        self._coef_chunk = None
        self._pot_key = None
        self._pot_obj = None
        self._pot_dev = None
        # with the default table height, scalar fields without a potential take the whole-step
        # form the autotuner measures fastest on this member table (LDS tiles or register row
        # strips; every form gives the same result); False keeps the LDS tiles
        self.autotune = True
        # rk4_cuda_solver.py:2280-2294: the two stencils, and the 9-point one needs dx == dy
        self.laplacian = getattr(config.solver, "laplacian", "five-point")
        if self.laplacian not in ("five-point", "isotropic-9pt"):
            raise ValueError(f"solver.laplacian='{self.laplacian}' is not recognised. "
                             "Valid values: 'five-point', 'isotropic-9pt'.")
        if self.laplacian == "isotropic-9pt":
            rel = abs(grid.dx - grid.dy) / max(grid.dx, grid.dy)
            if rel > 1e-8:
                raise ValueError(f"solver.laplacian='isotropic-9pt' requires dx == dy, got dx={grid.dx}, "
                                 f"dy={grid.dy} (relative difference {rel:.2e}).")
        self.boundary = getattr(grid, "boundary", "periodic")
        # the spinor stencil (rk4_cuda_solver.py:1600-1615) is the five-point Laplacian on periodic
        # or closed-interval grids; the isotropic 9-point stencil exists for scalar fields only
        if self.components == 2 and self.laplacian != "five-point":
            raise SolverCompatibilityError("rk4-fdm-hip runs spinor fields with the five-point stencil "
                                           "(solver.laplacian='isotropic-9pt' is for scalar fields)")

    def member_specs(self, model, batch):
        specs = super().member_specs(model, batch)
        for s in specs:
            s.laplacian = self.laplacian
            s.boundary = self.boundary
        return specs

    def _potential(self, model, batch):
        """Packed complex potential (shared or one per member) or None for V = 0."""
        pot = model.potential
        if pot is None or np.ndim(pot) == 0:
            if pot not in (None, 0, 0.0):
                raise ValueError("a constant nonzero potential must be given as an array")
            return None
        if self.components == 2:
            raise SolverCompatibilityError("rk4-fdm-hip takes no potential for spinor fields")
        pot = np.asarray(pot)
        key = (pot.shape, pot.dtype.str, batch)
        # the cache holds the host object itself (compared by identity, as the reference does,
        # ifrk_cuda_fused.py:643-654), so a new array can never reuse a stale device copy
        if self._pot_key != key or self._pot_obj is not model.potential:
            ny, nx = self.grid.shape
            arr = pot.reshape(-1, ny, nx) if pot.size % (ny * nx) == 0 else np.broadcast_to(pot, (ny, nx))[None]
            if arr.shape[0] not in (1, batch):
                raise ValueError(f"potential holds {arr.shape[0]} members for a batch of {batch}")
            self._pot_dev = DeviceArray.from_numpy(np.ascontiguousarray(arr, dtype=self.cdtype).reshape(-1))
            self._pot_shared = arr.shape[0] == 1
            self._pot_key = key
            self._pot_obj = model.potential
        return self._pot_dev

    def _ensure_engine(self, model, batch):
        pot = self._potential(model, batch)
        tab = self.ensure_table(model, batch)
        if pot is not None:
            per = 0 if self._pot_shared else self.grid.nx * self.grid.ny
            if any(s.pot_off != m * per for m, s in enumerate(tab.specs)):
                tab.set_pot_offsets([m * per for m in range(batch)])
        # one whole-step launch per step where it applies (scalar with no potential, or spinor
        # in single precision on a periodic grid); the four stage launches otherwise -- both give
        # the same result
        strip_h = None
        if self.components == 1:
            whole = pot is None and tab.tile_h in (16, 32, 64)
            if whole and tab.tile_h == 16 and self.autotune:
                strip_h = self._tuned_strip_height(tab)
        else:
            whole = spinor_whole_step_applies(tab, self.precision)
        e = self.engine
        if e is None or e.table is not tab or e.whole_step != whole or e.strip_h != strip_h:
            if e is not None and e.table is tab:
                self._carry_state(whole, strip_h)
            else:
                self.engine = Rk4Engine(tab, self._pump_dev, self.precision, self.reservoir_kind, self.stream,
                                        whole_step=whole, strip_h=strip_h)
            self._const_coef = None
        self.engine.potential = pot
        if getattr(self.engine, "_dt", None) != self.config.solver.dt:
            self.engine.set_dt(self.config.solver.dt)
        return self.engine

    def _tuned_strip_height(self, tab):
        """Strip height of the fastest whole-step form on this table of height 16 (None: LDS tiles),
        from the autotuner's cache or its timing of LDS tiles and every strip height."""
        if any(s.nx * s.ny > STRIP_MAX_POINTS for s in tab.specs):
            return None
        opts = [c for c in candidates(self.precision, 1) if c.whole_step and c.table_h == 16]
        choice = tune_rk4(tab.specs, self.precision, self.reservoir_kind, options=opts)
        return choice.strip_h

    def _carry_state(self, whole, strip_h=None):
        """Switch launch form on the same table, keeping the current field and reservoir."""
        old = self.engine
        self.engine = Rk4Engine(old.table, self._pump_dev, self.precision, self.reservoir_kind, self.stream,
                                whole_step=whole, strip_h=strip_h)
        self.engine.psi[0].copy_from(old.psi[0], self.stream)
        for a, b in zip(self.engine.res, old.res):
            a.copy_from(b, self.stream)
        self.stream.synchronize()

    def _adopt_state(self, model, state):
        e = self.engine
        f = state.field
        if f.ptr == e.psi[1].ptr:
            e.psi.reverse()
        elif f.ptr != e.psi[0].ptr:
            e.psi[0].copy_from(f.reshape(-1), self.stream)
        for i, field in enumerate(model.reservoir.get_state()):
            if field.ptr != e.res[i].ptr:
                e.res[i].copy_from(field.reshape(-1), self.stream)

    def _publish_state(self, model, state):
        e = self.engine
        state._field = e.psi[0].reshape(state._canonical_shape)
        state._field_k = None
        model.reservoir.assign_device(tuple(x.reshape(model.reservoir.shape) for x in e.res))

    def _coef_const(self, model, batch):
        if self._const_coef is None:
            vals = self.pump_coefficients(model, [0.0], batch)[0]
            self._const_coef = DeviceArray.from_numpy(np.tile(vals, (4, 1)).astype(self.rdtype))
        return self._const_coef

    def _coef_for_steps(self, model, batch, t0, first_step, count):
        dt = self.config.solver.dt
        times = []
        for i in range(count):
            t = t0 + (first_step + i) * dt
            times.extend((t, t + 0.5 * dt, t + 0.5 * dt, t + dt))
        vals = self.pump_coefficients(model, times, batch).astype(self.rdtype).reshape(-1)
        if self._coef_chunk is None or self._coef_chunk.size < vals.size:
            self._coef_chunk = DeviceArray.empty((vals.size,), self.rdtype)
        self._coef_chunk.view_range(0, (vals.size,)).upload(vals, self.stream)
        return self._coef_chunk

    def _run(self, model, state, n_steps, noise):
        batch = state.batch_size
        self.batch_physics.check(state._canonical_shape, state.components)
        e = self._ensure_engine(model, batch)
        self._adopt_state(model, state)
        const = self.pump_time_independent(model)
        use_noise = bool(noise and model.noise is not None and model.noise.active)
        graphs = bool(self.config.solver.cuda_graphs)
        done = 0
        while done < n_steps:
            chunk = min(n_steps - done, self.max_graph_chunk)
            first = state.step + done
            coef = self._coef_const(model, batch) if const else \
                self._coef_for_steps(model, batch, state.t0, first, chunk)
            e.run_steps(chunk, coef, const, use_noise, first, graphs)
            done += chunk
        self.stream.synchronize()
        state.step += n_steps
        state.t = state.t0 + state.step * self.config.solver.dt
        self._publish_state(model, state)

    def step(self, model, state, dt):
        """One RK4 step without noise (the reference adds noise in ``advance``)."""
        self._run(model, state, 1, noise=False)
        return ACCEPTED

    def advance(self, model, state, n_steps: int) -> None:
        if n_steps > 0:
            self._run(model, state, n_steps, noise=True)
