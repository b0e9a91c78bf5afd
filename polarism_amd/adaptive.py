"""Adaptive trial control: PI step sizes, tick landing, and where the decision runs.

``PIControl`` holds what the controller remembers between trials (the next proposal,
the previous error, whether growth is frozen after a rejection, and statistics).  The
same two operations run either on the host or on the device:

* ``propose(t, target)`` -- the trial's h: the proposal, cut down so that the trial lands
  exactly on ``target`` when it would overshoot it; an h below 16 eps max(1, |t|) is an
  underflow error (solver/step_controller.py:48-56);
* ``settle(...)``        -- after the trial: on acceptance the PI factor
  0.9 err^(-0.7/k) err_prev^(0.4/k) in [0.2, 5] (no growth right after a rejection), on
  rejection max(0.2, 0.9 err^(-1/k)) <= 1; k = min(p, p_hat) + 1
  (solver/step_controller.py:57-81).

``host_trials`` drives a solver's ``step`` trial by trial (the reference's loop,
abstract_solver.py:86-96).  Solvers with device kernels for the whole trial instead
load the state into a ``pgpe_adaptive_state`` (``to_device`` / ``from_device``) and let
the ``pgpe_adaptive_*`` kernels run the identical arithmetic in FP64 between their
trial launches, so no per-trial host round trip is needed.
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass

from . import _native as N


@dataclass
class PIControl:
    k: float                      # min(order, embedded order) + 1
    h_next: float                 # the next trial's proposal
    dt_max: float | None = None   # largest allowed h (solver.dt_max)
    err_prev: float = 1.0
    no_growth: bool = False
    accepted: int = 0
    rejected: int = 0
    dt_min: float = math.inf
    dt_max_used: float = 0.0
    dt_last: float = 0.0

    @classmethod
    def for_solver(cls, solver) -> "PIControl":
        from .tableaus import TABLEAUS
        sc = solver.config.solver
        tab = getattr(solver, "tableau", TABLEAUS[sc.tableau])
        limit = sc.dt_max
        return cls(k=float(min(tab.order, tab.embedded_order) + 1),
                   h_next=min(sc.dt, limit if limit is not None else math.inf), dt_max=limit)

    # -- the two operations -------------------------------------------------------
    def propose(self, t: float, target: float) -> tuple[float, bool, float]:
        """(h, clamped, proposal) of the next trial from time ``t`` towards ``target``."""
        proposal = self.h_next
        clamped = proposal >= target - t
        h = target - t if clamped else proposal
        if h < 16 * sys.float_info.epsilon * max(1.0, abs(t)):
            raise RuntimeError(f"adaptive step size underflow at t={t}")
        return h, clamped, proposal

    def settle(self, h: float, clamped: bool, proposal: float, accepted: bool, error,
               dt_next: float | None = None) -> None:
        """Update the memory after a trial with norm ``error`` (None / NaN / negative count as inf)."""
        err = math.inf if (error is None or not math.isfinite(error) or error < 0) else float(error)
        if accepted:
            self.accepted += 1
            self.dt_min = min(self.dt_min, h)
            self.dt_max_used = max(self.dt_max_used, h)
            self.dt_last = h
            grow = 5.0 if err == 0 else 0.9 * err ** (-0.7 / self.k) * self.err_prev ** (0.4 / self.k)
            grow = min(5.0, max(0.2, grow))
            factor = min(1.0, grow) if self.no_growth else grow
            self.err_prev = max(err, 1e-4)
            self.no_growth = False
        else:
            self.rejected += 1
            factor = min(1.0, max(0.2, 0.9 * err ** (-1 / self.k)))
            self.no_growth = True
        nxt = dt_next if dt_next is not None else h * factor
        if accepted and clamped:
            nxt = max(nxt, proposal)
        self.h_next = nxt if self.dt_max is None else min(nxt, self.dt_max)

    # -- device exchange ------------------------------------------------------------
    def to_device(self, t: float, target: float, members: int, count: float, blend: bool) -> N.AdaptiveState:
        st = N.AdaptiveState()
        st.t, st.target = float(t), float(target)
        st.h_next, st.err_prev = float(self.h_next), float(self.err_prev)
        st.dt_max = float(self.dt_max) if self.dt_max is not None else 0.0
        st.k = float(self.k)
        st.dt_min, st.dt_max_used, st.dt_last = float(self.dt_min), float(self.dt_max_used), float(self.dt_last)
        st.count = float(count)
        st.members, st.blend = int(members), int(bool(blend))
        st.no_growth = int(self.no_growth)
        st.accepted, st.rejected = int(self.accepted), int(self.rejected)
        st.done = int(not t < target)
        return st

    def from_device(self, st: N.AdaptiveState) -> None:
        if st.status == N.ADAPTIVE_UNDERFLOW:
            raise RuntimeError(f"adaptive step size underflow at t={st.t}")
        self.h_next, self.err_prev = st.h_next, st.err_prev
        self.no_growth = bool(st.no_growth)
        self.accepted, self.rejected = int(st.accepted), int(st.rejected)
        self.dt_min, self.dt_max_used, self.dt_last = st.dt_min, st.dt_max_used, st.dt_last

    # -- checkpoints (simulation_controller.py:736-923 stores these keys) ------------
    @property
    def statistics(self) -> dict:
        return dict(accepted=self.accepted, rejected=self.rejected, dt_min=self.dt_min,
                    dt_max=self.dt_max_used, dt_last=self.dt_last)

    def state_dict(self) -> dict:
        return dict(h_next=self.h_next, err_prev=self.err_prev, **self.statistics)

    def load_state_dict(self, values: dict) -> None:
        self.h_next = float(values["h_next"])
        self.err_prev = float(values["err_prev"])
        self.no_growth = False
        self.accepted, self.rejected = int(values["accepted"]), int(values["rejected"])
        self.dt_min, self.dt_max_used = float(values["dt_min"]), float(values["dt_max"])
        self.dt_last = float(values["dt_last"])


def host_trials(control: PIControl, solver, model, state, target: float) -> None:
    """Trials through ``solver.step`` until ``state.t`` reaches ``target``; an accepted trial
    that was cut to the target lands on it exactly."""
    while state.t < target:
        h, clamped, proposal = control.propose(state.t, target)
        report = solver.step(model, state, h)
        if report.accepted and clamped:
            state.t = target
        control.settle(h, clamped, proposal, report.accepted, report.error, report.dt_next)
