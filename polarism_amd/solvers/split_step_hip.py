"""Split-step FFT methods on gfx950.

``split-step-fft-hip`` is a drop-in for the reference's CPU ``split-step-fft``
(solver/split_step_fft_solver.py:23-159): Strang splitting
R(dt/2) K(dt/2) N(dt) noise K(dt/2) R(dt/2), four transforms per step, noise
drawn with the reservoir after the first half-step.

A spinor field (psi+, psi-) takes the 2x2 TE-TM / Zeeman propagator exp(-i H(k) dt / 2 hbar)
(dispersion.py:66-83) in the column passes, per-component or shared reservoirs and the
cross-spin phase alpha_2 |psi_o|^2 in the local step (split_step_fft_solver.py:46-53, 81-94,
141-143).

``split-step-fft-kouter-hip`` follows the order of the reference's
``split-step-fft-cuda`` (solver/split_step_fft_cuda_solver.py:106-184,
1144-1325): K(dt/2) [R(dt/2) N(dt) R(dt/2) noise] K(dt/2), noise with the
reservoir after the second half-step.

Each step is five fused launches (``pgpe_split_step``): rows, columns with
the kinetic propagator between a forward and an inverse transform, rows with
the local physics between an inverse and a forward transform, columns, rows.
Consecutive steps of the scalar CPU order are chained: the last row pass of a
step and the first of the next become one (the field stays in the scratch
buffer between them), four launches per step inside a chunk, bit-identical to
separate steps.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N
from ..device import DeviceArray, Graph
from ..solver_api import ACCEPTED, SolverCapabilities, register_solver
from .base import HipSolverBase

_ALL_RES = frozenset({"single", "double", "quadratic-double", "none"})


class FFTPlan:
    """Owning handle of a ``pgpe_fft_plan``."""

    def __init__(self, nx, ny, dx, dy, precision):
        p = C.c_void_p()
        N.call("pgpe_fft_plan_create", nx, ny, dx, dy, 0 if precision == "single" else 1, C.byref(p))
        self.handle = p.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and N._lib is not None:
            N._lib.pgpe_fft_plan_destroy(h)
            self.handle = None


class _SplitStepBase(HipSolverBase):
    ordering = 0
    injects_noise = True
    max_graph_chunk = 1000

    def __init__(self, config, grid, model):
        super().__init__(config, grid, model)
        self.res_components = getattr(model.reservoir, "components", 1)
        self.plan = FFTPlan(grid.nx, grid.ny, grid.dx, grid.dy, self.precision)
        self._psi = None
        self._tmp = None
        self._res = None
        self._const = None
        self._coef_chunk = None
        self._step_base = DeviceArray.zeros((1,), np.uint64)
        self._args = N.SplitArgs()

    def _ensure(self, model, batch):
        tab = self.ensure_table(model, batch)
        n = tab.total_points * self.components
        if self._psi is None or self._psi.size != n:
            self._psi = DeviceArray.empty((n,), self.cdtype)
            self._tmp = DeviceArray.empty((n,), self.cdtype)
            nr = tab.total_points * self.res_components
            self._res = [DeviceArray.empty((nr,), self.rdtype) for _ in range(self.nres)]
            self._graphs.clear()
            self._const = None
        return tab

    def _coefs(self, model, batch, t0, first, count, const):
        dt = self.config.solver.dt
        if const:
            if self._const is None:
                v = self.pump_coefficients(model, [0.0], batch)[0]
                # three rows: a chained step also reads the next step's first coefficient
                self._const = DeviceArray.from_numpy(np.tile(v, (3, 1)).astype(self.rdtype))
            return self._const
        times = []
        for i in range(count):
            t = t0 + (first + i) * dt
            times.extend((t + dt / 4, t + 3 * dt / 4))
        vals = self.pump_coefficients(model, times, batch).astype(self.rdtype).reshape(-1)
        if self._coef_chunk is None or self._coef_chunk.size < vals.size:
            self._coef_chunk = DeviceArray.empty((vals.size,), self.rdtype)
        self._coef_chunk.view_range(0, (vals.size,)).upload(vals, self.stream)
        return self._coef_chunk

    def _chain(self, k, chunk):
        """Chain bits of step k of a chunk: the scalar CPU order keeps the field in the scratch
        buffer between consecutive steps; the chunk's last step writes the state."""
        if self.ordering != 0 or self.components != 1:
            return 0
        return (N.CHAIN_IN if k > 0 else 0) | (N.CHAIN_OUT if k < chunk - 1 else 0)

    def _launch(self, tab, coef_ptr, noise, step, step_base, chain=0):
        a = self._args
        a.precision = self.prec_code
        a.reservoir = N.RESERVOIR_KIND[self.reservoir_kind]
        a.ordering = self.ordering
        a.noise = 1 if noise else 0
        a.members = tab.members
        a.components = self.components
        a.res_components = self.res_components
        a.dt = float(self.config.solver.dt)
        a.table = tab.table_ptr
        a.plan = self.plan.handle
        a.psi = self._psi.ptr
        a.tmp = self._tmp.ptr
        for f in range(self.nres):
            a.res[f] = self._res[f].ptr
        a.pump = self._pump_dev.ptr
        a.coef = coef_ptr
        a.step = int(step)
        a.step_base = step_base
        a.chain = chain
        N.check(N.lib().pgpe_split_step(C.byref(a), self.stream.handle), "pgpe_split_step")

    def _run(self, model, state, n_steps):
        batch = state.batch_size
        self.batch_physics.check(state._canonical_shape, state.components)
        tab = self._ensure(model, batch)
        f = state.field
        if f.ptr != self._psi.ptr:
            self._psi.copy_from(f.reshape(-1), self.stream)
        for i, r in enumerate(model.reservoir.get_state()):
            if r.ptr != self._res[i].ptr:
                self._res[i].copy_from(r.reshape(-1), self.stream)
        const = self.pump_time_independent(model)
        noise = model.noise is not None and model.noise.active
        item = np.dtype(self.rdtype).itemsize
        graphs = bool(self.config.solver.cuda_graphs)
        done = 0
        while done < n_steps:
            chunk = min(n_steps - done, self.max_graph_chunk)
            first = state.step + done
            coef = self._coefs(model, batch, state.t0, first, chunk, const)
            if graphs and chunk > 1:
                key = (chunk, const, noise, coef.ptr, self._psi.ptr)
                g = self._graphs.get(key)
                if g is None:
                    def body(chunk=chunk, coef=coef):
                        for k in range(chunk):
                            ptr = coef.ptr if const else coef.ptr + k * 2 * batch * item
                            self._launch(tab, ptr, noise, k, self._step_base.ptr, self._chain(k, chunk))
                    self.stream.synchronize()
                    g = Graph.capture(self.stream, body)
                    if len(self._graphs) >= 4:
                        self._graphs.pop(next(iter(self._graphs)))
                    self._graphs[key] = g
                self._step_base.upload(np.asarray([first], dtype=np.uint64), self.stream)
                g.launch(self.stream)
            else:
                for k in range(chunk):
                    ptr = coef.ptr if const else coef.ptr + k * 2 * batch * item
                    self._launch(tab, ptr, noise, first + k, None, self._chain(k, chunk))
            done += chunk
        self.stream.synchronize()
        state.step += n_steps
        state.t = state.t0 + state.step * self.config.solver.dt
        state._field = self._psi.reshape(state._canonical_shape)
        state._field_k = None
        model.reservoir.assign_device(tuple(x.reshape(model.reservoir.shape) for x in self._res))

    def step(self, model, state, dt):
        self._run(model, state, 1)
        return ACCEPTED

    def advance(self, model, state, n_steps: int) -> None:
        if n_steps > 0:
            self._run(model, state, n_steps)


_CAPS = dict(components=frozenset({1, 2}), ndim=frozenset({2}), grid_types=frozenset({"periodic"}),
             reservoirs=_ALL_RES, custom_reservoirs="unsupported", potential=False, kinetic_relaxation=False,
             dispersion=frozenset({"parabolic"}), terms=False, noise=frozenset({"none", "wigner"}),
             time_dependent_potential=False, quantitative_reservoirs=_ALL_RES)


@register_solver("split-step-fft-hip")
class SplitStepFFTHipSolver(_SplitStepBase):
    capabilities = SolverCapabilities(
        **_CAPS, description="Strang split-step FFT on gfx950 in the reference CPU order "
                             "(R K N noise K R): LDS Stockham transforms with the propagator and the "
                             "local physics fused in, five launches per step, four when chained.")
    ordering = 0


@register_solver("split-step-fft-kouter-hip")
class SplitStepFFTKOuterHipSolver(_SplitStepBase):
    capabilities = SolverCapabilities(
        **_CAPS, description="Strang split-step FFT on gfx950 in the K-outer order of the reference's "
                             "split-step-fft-cuda (K [R N R noise] K).")
    ordering = 1
