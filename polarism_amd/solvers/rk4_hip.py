"""``rk4-fdm-hip``: the stage-coupled RK4 finite-difference stepper on gfx950.

Drop-in for the reference's ``rk4-fdm`` / ``rk4-cuda`` methods
(solver/rk4_fdm_solver.py:14-77, solver/rk4_cuda_solver.py:2197-3642):
the same RK4 stages, reservoir clamp and noise semantics, with each stage one
fused launch of ``pgpe_rk4_stage`` over every member of the batch.  Chunks
of steps are captured as HIP graphs whose pump coefficients and noise step
base live in device buffers, so replaying a chunk needs no host sync.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _native as N
from ..device import DeviceArray
from ..solver_api import ACCEPTED, SolverCapabilities, register_solver
from .base import HipSolverBase

_ALL_RES = frozenset({"single", "double", "quadratic-double", "none"})


@register_solver("rk4-fdm-hip")
class RK4FDMHipSolver(HipSolverBase):
    capabilities = SolverCapabilities(
        components=frozenset({1}), ndim=frozenset({2}), grid_types=frozenset({"periodic"}),
        reservoirs=_ALL_RES, custom_reservoirs="unsupported", potential=False,
        kinetic_relaxation=True, dispersion=frozenset({"parabolic"}), terms=False,
        noise=frozenset({"none", "wigner"}), quantitative_reservoirs=_ALL_RES,
        description="Stage-coupled FDM RK4 on gfx950: one fused HIP launch per RK stage over all "
                    "members, LDS halo tiles, in-kernel Wigner noise, HIP-graph chunks. Periodic 2-D, "
                    "zero potential.")
    injects_noise = True   # advance() adds the Wigner increment inside the final stage kernel
    max_graph_chunk = 1000

    def __init__(self, config, grid, model):
        super().__init__(config, grid, model)
        if model.components != 1:
            raise ValueError("rk4-fdm-hip is scalar; use rk4-fdm-spinor-hip for spinors")
        self._bufs = None
        self._step_base = DeviceArray.zeros((1,), np.uint64)
        self._const_coef = None
        self._coef_chunk = None

    # -- buffers -------------------------------------------------------------
    def _ensure_buffers(self, model, batch):
        tab = self.ensure_table(model, batch)
        n = tab.total_points
        if self._bufs is None or self._bufs["n"] != n:
            c, r = self.cdtype, self.rdtype
            self._bufs = dict(
                n=n,
                psi=[DeviceArray.empty((n,), c), DeviceArray.empty((n,), c)],
                k=[DeviceArray.empty((n,), c) for _ in range(3)],
                kres=[[DeviceArray.empty((n,), r) for _ in range(self.nres)] for _ in range(3)],
                res=[DeviceArray.empty((n,), r) for _ in range(self.nres)],
            )
            self._graphs.clear()
            self._const_coef = None
        return tab

    def _adopt_state(self, model, state):
        """Make the solver-owned buffers hold the current state (device copies only)."""
        b = self._bufs
        f = state.field
        if f.ptr == b["psi"][1].ptr:
            b["psi"].reverse()
        elif f.ptr != b["psi"][0].ptr:
            b["psi"][0].copy_from(f.reshape(-1), self.stream)
        res = model.reservoir.get_state()
        for i, field in enumerate(res):
            if field.ptr != b["res"][i].ptr:
                b["res"][i].copy_from(field.reshape(-1), self.stream)

    def _publish_state(self, model, state):
        b = self._bufs
        shape = state._canonical_shape
        state._field = b["psi"][0].reshape(shape)
        state._field_k = None
        rshape = model.reservoir.shape
        model.reservoir.assign_device(tuple(x.reshape(rshape) for x in b["res"]))

    # -- coefficients ----------------------------------------------------------
    def _stage_times(self, t, dt):
        return (t, t + 0.5 * dt, t + 0.5 * dt, t + dt)

    def _coef_const(self, model, batch):
        if self._const_coef is None:
            vals = self.pump_coefficients(model, [0.0], batch)[0]
            arr = np.tile(vals, (4, 1)).astype(self.rdtype)
            self._const_coef = DeviceArray.from_numpy(arr)
        return self._const_coef

    def _coef_for_steps(self, model, batch, t0_step, count):
        """(count, 4, batch) coefficients of the steps starting at tick t0_step."""
        dt = self.config.solver.dt
        times = []
        for i in range(count):
            t = self._tick_time(t0_step + i)
            times.extend(self._stage_times(t, dt))
        vals = self.pump_coefficients(model, times, batch).reshape(count, 4, batch).astype(self.rdtype)
        if self._coef_chunk is None or self._coef_chunk.size < vals.size:
            self._coef_chunk = DeviceArray.empty((max(vals.size, 4 * batch),), self.rdtype)
        self._coef_chunk.view_range(0, (vals.size,)).upload(vals.reshape(-1), self.stream)
        return self._coef_chunk

    def _tick_time(self, step):
        return self._t0 + step * self.config.solver.dt

    # -- launches --------------------------------------------------------------
    def _launch_step(self, tab, psi_in, psi_out, coef_ptrs, noise, step, step_base_ptr):
        b = self._bufs
        a = N.Rk4Args()
        a.precision = self.prec_code
        a.reservoir = N.RESERVOIR_KIND[self.reservoir_kind]
        a.num_tiles = tab.num_tiles
        a.tile_y = tab.tile_h
        a.members = tab.members
        a.dt = float(self.config.solver.dt)
        a.table = tab.table_ptr
        a.tile_member = tab.tile_member_ptr
        a.psi0 = psi_in.ptr
        a.pump = self._pump_dev.ptr
        for f in range(self.nres):
            a.res0[f] = b["res"][f].ptr
        for s in range(4):
            a.stage = s
            a.coef = coef_ptrs[s]
            if s > 0:
                a.kprev = b["k"][s - 1].ptr
                for f in range(self.nres):
                    a.kres_prev[f] = b["kres"][s - 1][f].ptr
            if s < 3:
                a.kout = b["k"][s].ptr
                for f in range(self.nres):
                    a.kres_out[f] = b["kres"][s][f].ptr
            else:
                a.k1, a.k2 = b["k"][0].ptr, b["k"][1].ptr
                a.psi_out = psi_out.ptr
                for f in range(self.nres):
                    a.kres1[f] = b["kres"][0][f].ptr
                    a.kres2[f] = b["kres"][1][f].ptr
                    a.res_out[f] = b["res"][f].ptr
                a.noise = 1 if noise else 0
                a.step = int(step)
                a.step_base = step_base_ptr
            N.check(N.lib().pgpe_rk4_stage(C.byref(a), self.stream.handle), "pgpe_rk4_stage")

    def _coef_ptrs(self, coef, index, batch):
        item = np.dtype(self.rdtype).itemsize
        return [coef.ptr + ((index * 4 + s) * batch) * item for s in range(4)]

    # -- solver contract ---------------------------------------------------------
    def _run(self, model, state, n_steps, noise):
        batch = state.batch_size
        self.batch_physics.check(state._canonical_shape, state.components)
        tab = self._ensure_buffers(model, batch)
        self._t0 = state.t0
        self._adopt_state(model, state)
        b = self._bufs
        const = self.pump_time_independent(model)
        use_noise = bool(noise and model.noise is not None and model.noise.active)
        graphs = bool(self.config.solver.cuda_graphs)
        done = 0
        while done < n_steps:
            remaining = n_steps - done
            step0 = state.step + done
            chunk = min(remaining, self.max_graph_chunk)
            if graphs and chunk >= 2:
                chunk -= chunk % 2
                coef = self._coef_const(model, batch) if const else self._coef_for_steps(model, batch, step0, chunk)
                p0, p1 = b["psi"]
                key = (chunk, const, use_noise, tab.tile_h, p0.ptr, coef.ptr)

                def body(chunk=chunk, coef=coef):
                    for k in range(chunk):
                        src, dst = (p0, p1) if k % 2 == 0 else (p1, p0)
                        ptrs = self._coef_ptrs(coef, 0 if const else k, batch)
                        self._launch_step(tab, src, dst, ptrs, use_noise, k, self._step_base.ptr)

                graph = self.capture(key, body)
                self._step_base.upload(np.asarray([step0], dtype=np.uint64), self.stream)
                graph.launch(self.stream)
            else:
                chunk = 1
                coef = self._coef_const(model, batch) if const else self._coef_for_steps(model, batch, step0, 1)
                self._launch_step(tab, b["psi"][0], b["psi"][1], self._coef_ptrs(coef, 0, batch),
                                  use_noise, step0, None)
                b["psi"].reverse()
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
        if n_steps <= 0:
            return
        self._run(model, state, n_steps, noise=True)
