"""``ifrk-fft-hip`` / ``ifrk4-fft-hip`` / ``ip-rk4-hip``: Lawson integrating-factor RK on gfx950.

Drop-in for the reference's ``ifrk-fft-cuda`` / ``ifrk4-fft-cuda``
(solver/ifrk_fft_cuda_solver.py:24-336): stage-coupled reservoir, RK4, DP5 or
DOP853 tableaus, fixed steps or adaptive trials with the embedded error
(DP5: sqrt(err5 / count); DOP853: err5 / sqrt((err5 + 0.01 err3) count); max
over members) and the PI controller; scalar or spinor (TE-TM + Zeeman 2x2
exponential), optional complex potential and kinetic relaxation.

Each stage is two or three fused launches of ``pgpe_lawson`` (combination +
inverse transform along y; inverse along x + RHS + forward along x; forward
along y); the error norm is reduced per member on the device in double and
only 16 bytes per member return to the host per trial.
"""
from __future__ import annotations

import ctypes as C
import math
from collections import deque

import numpy as np

from .. import _native as N
from ..device import DeviceArray, Event, PinnedBuffer
from ..solver_api import ACCEPTED, SolverCapabilities, StepReport, register_solver
from ..tableaus import TABLEAUS
from .base import HipSolverBase
from .split_step_hip import FFTPlan

_ALL_RES = frozenset({"single", "double", "quadratic-double", "none"})


@register_solver("ifrk-fft-hip")
class IFRKFFTHipSolver(HipSolverBase):
    capabilities = SolverCapabilities(
        components=frozenset({1, 2}), ndim=frozenset({2}), grid_types=frozenset({"periodic"}),
        reservoirs=_ALL_RES, custom_reservoirs="unsupported", potential=True, kinetic_relaxation=True,
        dispersion=frozenset({"parabolic"}), terms=False, noise=frozenset({"none", "wigner"}), adaptive=True,
        quantitative_reservoirs=_ALL_RES,
        description="Lawson interaction-picture RK (rk4 / dp5 / dop853, fixed or adaptive) on gfx950 with "
                    "LDS Stockham transforms and fused stage kernels; scalar or TE-TM/Zeeman spinor.")

    def __init__(self, config, grid, model):
        super().__init__(config, grid, model)
        self.tableau = TABLEAUS[getattr(self, "tableau_name", config.solver.tableau)]
        A, b, c, b_hat, b_hat3 = self.tableau.arrays()
        self._A, self._b, self._c = A, b, c
        self._w5 = None if b_hat is None else b - b_hat
        self._w3 = None if b_hat3 is None else b - b_hat3
        self.plan = FFTPlan(grid.nx, grid.ny, grid.dx, grid.dy, self.precision)
        self.res_components = getattr(model.reservoir, "components", 1)
        self.lap = 1 if self.batch_physics.any_nonzero("kinetic_relaxation_eta") else 0
        self._bufs = None
        self._pot_key = None
        self._pot_obj = None
        self._pot_dev = None
        self._args = N.LawsonArgs()
        self._k_valid = False
        self._coef_key = None   # (buffer, values) of the last coefficient upload
        self._ctl_dev = None    # device pgpe_adaptive_state of the device-controlled trials
        self._ctl_host = None   # pinned copies of it, one per queued trial
        self._ctl_events = None

    # -- buffers -------------------------------------------------------------
    def _ensure(self, model, batch):
        tab = self.ensure_table(model, batch)
        C_ = self.components
        n = tab.total_points * C_
        nr = tab.total_points * self.res_components
        s = self.tableau.stages
        if self._bufs is None or self._bufs["n"] != n:
            c, r = self.cdtype, self.rdtype
            self._coef_key = None   # new buffers hold no coefficients yet
            self._bufs = dict(
                n=n,
                psi=DeviceArray.empty((n,), c), psi_new=DeviceArray.empty((n,), c),
                psik=DeviceArray.empty((n,), c), psik_new=DeviceArray.empty((n,), c),
                work=DeviceArray.empty((n,), c), aux=DeviceArray.empty((2 * n,), c),
                f=[DeviceArray.empty((n,), c) for _ in range(s)],
                dres=[[DeviceArray.empty((nr,), r) for _ in range(self.nres)] for _ in range(s)],
                res=[DeviceArray.empty((nr,), r) for _ in range(self.nres)],
                res_new=[DeviceArray.empty((nr,), r) for _ in range(self.nres)],
                coef=DeviceArray.empty((s * batch,), r),
                partials=DeviceArray.zeros((2 * batch * self.grid.ny,), np.float64),
                sums=DeviceArray.zeros((2 * batch,), np.float64),
            )
            self._k_valid = False
        return tab

    def _potential_ptr(self, model):
        pot = model.potential
        if pot is None or np.ndim(pot) == 0:
            if pot not in (None, 0, 0.0):
                raise ValueError("a constant nonzero potential must be given as an array")
            return None
        pot = np.asarray(pot)
        key = (pot.shape, pot.dtype.str)
        # the cache holds the host object itself (compared by identity, as the reference does,
        # ifrk_cuda_fused.py:643-654), so a new array can never reuse a stale device copy
        if self._pot_key != key or self._pot_obj is not model.potential:
            arr = np.broadcast_to(pot, self.grid.shape).astype(self.cdtype)
            self._pot_dev = DeviceArray.from_numpy(arr)
            self._pot_key = key
            self._pot_obj = model.potential
        return self._pot_dev.ptr

    def _adopt(self, model, state):
        b = self._bufs
        f = state.field
        if f.ptr != b["psi"].ptr:
            b["psi"].copy_from(f.reshape(-1), self.stream)
            self._k_valid = False
            fk = state.field_k
            if fk is not None:  # a restored spectral cache: continue from it, not from FFT(psi)
                b["psik"].copy_from(fk.reshape(-1), self.stream)
                self._k_valid = True
        for i, r in enumerate(model.reservoir.get_state()):
            if r.ptr != b["res"][i].ptr:
                b["res"][i].copy_from(r.reshape(-1), self.stream)

    def _publish(self, model, state):
        """The real-space field, and its spectral cache when the solver holds psi^k of the
        same state (state.field_k, simulation_state.py:92-192; a checkpoint saves it)."""
        b = self._bufs
        state._field = b["psi"].reshape(state._canonical_shape)
        state._field_k = b["psik"].reshape(state._canonical_shape) if self._k_valid else None
        model.reservoir.assign_device(tuple(x.reshape(model.reservoir.shape) for x in b["res"]))

    # -- one trial ------------------------------------------------------------
    def _upload_coefficients(self, model, t, h, batch):
        """Pump coefficients of the stage times t + c_i h.  The upload synchronises the stream,
        so it is skipped when the values are already there (time-independent pumps) and fixed
        steps queue back to back."""
        b = self._bufs
        coefs = self.pump_coefficients(model, [t + float(ci) * h for ci in self._c], batch).astype(self.rdtype)
        key = (b["coef"].ptr, coefs.tobytes())
        if key != self._coef_key:
            b["coef"].upload(coefs.reshape(-1), self.stream)
            self._coef_key = key

    def _queue_trial(self, model, tab, h, adaptive, noise, step, ctl_ptr=None):
        """Queue the stage and final launches of one trial on the solver stream.

        With ``ctl_ptr`` (a device ``pgpe_adaptive_state``) the kernels read h from it and skip
        once it is done; the coefficients passed are then the bare tableau weights."""
        b = self._bufs
        batch = tab.members
        stages = self.tableau.stages
        A, bw, c = self._A, self._b, self._c
        pot = self._potential_ptr(model)
        item = np.dtype(self.rdtype).itemsize
        hw = 1.0 if ctl_ptr is not None else float(h)     # weight scale folded in on the host
        a = self._args
        a.precision = self.prec_code
        a.reservoir = N.RESERVOIR_KIND[self.reservoir_kind]
        a.components = self.components
        a.res_components = self.res_components
        a.members = batch
        a.lap = self.lap
        a.h = float(h)
        a.atol = float(self.config.solver.atol)
        a.rtol = float(self.config.solver.rtol)
        a.table = tab.table_ptr
        a.plan = self.plan.handle
        a.psi0k = b["psik"].ptr
        a.work = b["work"].ptr
        a.aux = b["aux"].ptr
        a.psi_real = b["psi"].ptr
        a.pump = self._pump_dev.ptr
        a.potential = pot
        a.h_dev = None if ctl_ptr is None else ctl_ptr + N.AdaptiveState.h.offset
        a.skip_dev = None if ctl_ptr is None else ctl_ptr + N.AdaptiveState.done.offset
        for f in range(self.nres):
            a.res0[f] = b["res"][f].ptr
        for j in range(stages):
            a.f[j] = b["f"][j].ptr
            for f in range(self.nres):
                a.dres[j][f] = b["dres"][j][f].ptr
        lib = N.lib()
        # stages
        a.op = 0
        a.nerr = 0
        for i in range(stages):
            ci = float(c[i])
            a.frac0 = ci
            a.nterms = i
            for j in range(i):
                a.hc[j] = hw * float(A[i, j])
                a.fr[j] = ci - float(c[j])
                a.rc[j] = hw * float(A[i, j])
            a.stage_real = 1 if (i == 0 and not self.lap and self._k_valid is False) else 0
            a.coef = b["coef"].ptr + i * batch * item
            a.f_out = b["f"][i].ptr
            for f in range(self.nres):
                a.dres_out[f] = b["dres"][i][f].ptr
            N.check(lib.pgpe_lawson(C.byref(a), self.stream.handle), "pgpe_lawson stage")
        # final combination
        a.op = 1
        a.frac0 = 1.0
        a.nterms = stages
        for j in range(stages):
            a.hc[j] = hw * float(bw[j])
            a.fr[j] = 1.0 - float(c[j])
            a.rc[j] = hw * float(bw[j])
            a.he5[j] = hw * float(self._w5[j]) if (adaptive and self._w5 is not None) else 0.0
            a.he3[j] = hw * float(self._w3[j]) if (adaptive and self._w3 is not None) else 0.0
        a.nerr = 0 if not adaptive else (2 if self._w3 is not None else 1)
        a.coef = None
        a.noise = 1 if noise else 0
        a.step = int(step)
        a.step_base = None
        if adaptive:
            a.psi_out = b["psi_new"].ptr
            a.psik_out = b["psik_new"].ptr
            for f in range(self.nres):
                a.res_out[f] = b["res_new"][f].ptr
            a.psi_start = b["psi"].ptr
            a.partials = b["partials"].ptr
            a.error_sums = b["sums"].ptr
        else:
            a.psi_out = b["psi"].ptr
            a.psik_out = None if noise else b["psik"].ptr
            for f in range(self.nres):
                a.res_out[f] = b["res"][f].ptr
            a.psi_start = None
            a.partials = None
            a.error_sums = None
        N.check(lib.pgpe_lawson(C.byref(a), self.stream.handle), "pgpe_lawson final")

    def _error_count(self):
        """Points per member in the embedded error norm: field and reservoir values."""
        return self.components * self.grid.nx * self.grid.ny + self.nres * self.res_components * \
            self.grid.nx * self.grid.ny

    def _trial(self, model, state, h, adaptive, noise, step):
        tab = self._ensure(model, state.batch_size)
        b = self._bufs
        batch = tab.members
        self._upload_coefficients(model, state.t, h, batch)
        if not self._k_valid:
            N.call("pgpe_fft2", self.plan.handle, b["psi"].ptr, b["psik"].ptr, batch * self.components, 0,
                   self.stream.handle)
        self._queue_trial(model, tab, h, adaptive, noise, step)
        if not adaptive:
            self._k_valid = not noise
            return None
        sums = b["sums"].get(self.stream).reshape(batch, 2)
        count = self._error_count()
        err5, err3 = sums[:, 0], sums[:, 1]
        if self._w3 is not None:
            den = err5 + 0.01 * err3
            with np.errstate(divide="ignore", invalid="ignore"):
                member = np.where(den > 0, err5 / np.sqrt(np.where(den > 0, den, 1) * count), 0.0)
        else:
            member = np.sqrt(err5 / count)
        return float(np.max(member))

    # -- adaptive trials controlled on the device ----------------------------------------
    TRIALS_PER_GRAPH = 1
    device_trials = True      # False: decide every trial on the host (the reference's loop)

    def run_trials(self, model, state, target):
        """Trials until ``state.t`` reaches ``target``.  With a time-independent pump the whole
        loop runs on the device: a captured graph of ``TRIALS_PER_GRAPH`` trials, each bracketed
        by the controller kernels (propose, decide, commit), is replayed until the controller
        reports the target reached; the host reads the 144-byte state once per replay.  Pumps
        whose coefficient depends on the trial time take the host loop."""
        if not state.t < target:
            return
        if not (self.device_trials and self.pump_time_independent(model)):
            return super().run_trials(model, state, target)
        ctl = self.adaptive_control()
        batch = state.batch_size
        self.batch_physics.check(state._canonical_shape, state.components)
        tab = self._ensure(model, batch)
        self._adopt(model, state)
        b = self._bufs
        self._upload_coefficients(model, state.t, 0.0, batch)
        if not self._k_valid:
            N.call("pgpe_fft2", self.plan.handle, b["psi"].ptr, b["psik"].ptr, batch * self.components, 0,
                   self.stream.handle)
            self._k_valid = True
        if self._ctl_dev is None:
            self._ctl_dev = DeviceArray.empty((C.sizeof(N.AdaptiveState),), np.uint8)
        st = ctl.to_device(state.t, target, batch, self._error_count(), self._w3 is not None)
        self._ctl_dev.upload(np.frombuffer(bytes(st), dtype=np.uint8), self.stream)
        ptr = self._ctl_dev.ptr
        commit = N.CommitList()
        regions = [(b["psi"], b["psi_new"]), (b["psik"], b["psik_new"])]
        regions += [(b["res"][f], b["res_new"][f]) for f in range(self.nres)]
        for r, (dst, src) in enumerate(regions):
            commit.dst[r], commit.src[r], commit.bytes[r] = dst.ptr, src.ptr, dst.nbytes
        commit.count = len(regions)
        pot = self._potential_ptr(model)
        key = ("trials", self.TRIALS_PER_GRAPH, ptr, b["psi"].ptr, b["coef"].ptr, pot)

        def body():
            for _ in range(self.TRIALS_PER_GRAPH):
                N.call("pgpe_adaptive_propose", ptr, self.stream.handle)
                self._queue_trial(model, tab, 0.0, True, False, state.step, ctl_ptr=ptr)
                N.call("pgpe_adaptive_decide", ptr, b["sums"].ptr, self.stream.handle)
                N.check(N.lib().pgpe_adaptive_commit(ptr, C.byref(commit), self.stream.handle),
                        "pgpe_adaptive_commit")

        graph = self.capture(key, body)
        st = self._replay_trials(graph, st)
        ctl.from_device(st)
        state.t = st.t
        self._publish(model, state)

    SPECULATION = 2   # trial graphs kept queued ahead of the controller state the host has read

    def _replay_trials(self, graph, st):
        """Replay the one-trial graph until the device controller is done.  The host keeps up to
        ``SPECULATION`` trials queued (as many as the controller's last state says are still
        needed), each followed by an asynchronous copy of the state into pinned memory, so the
        GPU does not wait for the host between trials; a trial queued past the target finds
        the state done and its kernels return at once."""
        size = C.sizeof(N.AdaptiveState)
        if self._ctl_host is None:
            self._ctl_host = PinnedBuffer(self.SPECULATION * size)
            self._ctl_events = [Event() for _ in range(self.SPECULATION)]

        def expected(x):
            if x.done or x.status:
                return 0
            return max(1, math.ceil((x.target - x.t) / x.h_next))

        pending = deque()
        slot = 0
        want = expected(st)
        while want:
            while len(pending) < min(self.SPECULATION, want):
                graph.launch(self.stream)
                self._ctl_host.fetch(self._ctl_dev, slot * size, self.stream)
                self._ctl_events[slot].record(self.stream)
                pending.append(slot)
                slot = (slot + 1) % self.SPECULATION
            k = pending.popleft()
            self._ctl_events[k].synchronize()
            st = N.AdaptiveState.from_buffer_copy(self._ctl_host.read(k * size, size))
            want = expected(st)
        self.stream.synchronize()
        return st

    # -- solver contract ---------------------------------------------------------
    def step(self, model, state, dt):
        batch = state.batch_size
        self.batch_physics.check(state._canonical_shape, state.components)
        self._ensure(model, batch)
        self._adopt(model, state)
        if self.config.solver.adaptive:
            error = self._trial(model, state, dt, True, False, state.step)
            accepted = error <= 1
            if accepted:
                b = self._bufs
                b["psi"], b["psi_new"] = b["psi_new"], b["psi"]
                b["psik"], b["psik_new"] = b["psik_new"], b["psik"]
                b["res"], b["res_new"] = b["res_new"], b["res"]
                self._k_valid = True
                state.t += dt
            self.stream.synchronize()
            self._publish(model, state)
            return StepReport(accepted, error, None)
        self._trial(model, state, self.config.solver.dt, False, False, state.step)
        self.stream.synchronize()
        state.step += 1
        state.t = state.t0 + state.step * self.config.solver.dt
        self._publish(model, state)
        return ACCEPTED

    def advance(self, model, state, n_steps: int) -> None:
        if self.config.solver.adaptive:
            return super().advance(model, state, n_steps)
        if n_steps <= 0:
            return
        batch = state.batch_size
        self.batch_physics.check(state._canonical_shape, state.components)
        self._ensure(model, batch)
        self._adopt(model, state)
        noise = model.noise is not None and model.noise.active
        for _ in range(n_steps):
            self._trial(model, state, self.config.solver.dt, False, noise, state.step)
            state.step += 1
            state.t = state.t0 + state.step * self.config.solver.dt
        self.stream.synchronize()
        self._publish(model, state)


@register_solver("ifrk4-fft-hip")
class IFRK4FFTHipSolver(IFRKFFTHipSolver):
    """The RK4 tableau regardless of ``solver.tableau`` (ifrk_fft_cuda_solver.py:332-336)."""
    capabilities = SolverCapabilities(**{**IFRKFFTHipSolver.capabilities.__dict__, "adaptive": False})
    tableau_name = "rk4"


@register_solver("ip-rk4-hip")
class IPRK4HipSolver(IFRK4FFTHipSolver):
    """Drop-in for ``ip-rk4`` (solver/ip_rk4_solver.py:23-213).  The interaction-picture RK4
    carries u = exp(-L s) psi^k and applies exp(+-L dt/2), exp(+-L dt) around each stage; that is
    Lawson's RK4 (stage i: exp(L c_i h) psi0^k + h sum_j a_ij exp(L (c_i - c_j) h) f_j) with the
    products of the factors taken in closed form, so it runs on the Lawson RK4 kernels: same
    stage-coupled reservoir, spectral kinetic relaxation, scalar or TE-TM / Zeeman spinor.
    The reference's coherent drive, extra terms and time-dependent potentials are not taken."""
    capabilities = SolverCapabilities(**{
        **IFRK4FFTHipSolver.capabilities.__dict__,
        "description": "Interaction-picture RK4 (ip-rk4) on the Lawson RK4 kernels: stage-coupled reservoir, "
                       "FFT-only, periodic grid, scalar or TE-TM/Zeeman spinor."})
