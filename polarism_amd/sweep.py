"""Packed sweeps: many independent members, uneven grids, one launch per stage.

``PackedSweep`` runs every member of a parameter sweep in the same RK4 stage
launches over a packed member table (grids of different sizes included),
checks the steady-state criterion of the reference
(simulation_controller.py:710-734: relative L2 density change and relative
norm change both below ``rtol`` between checks ``check_every`` apart) on the
device, and retires converged members from the tile map so later launches
only touch members that are still running.  Members share the reservoir
model, precision and time step (the reference groups members that differ in
these into separate batches, sweep_config.py:18).

The reference runs each grid shape as a separate batch (sweep.py:128-138);
running all shapes in one launch is an addition of this package.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .device import DeviceArray, Stream, init
from .engine import Rk4Engine, complex_dtype, real_dtype
from .layout import MemberSpec, MemberTable
from .noise import make_initial_psi


@dataclass
class SweepMember:
    nx: int
    ny: int
    lx: float
    ly: float
    physics: object                 # PhysicsConstants-like (attribute access)
    envelope: np.ndarray            # (ny, nx) pump profile; the pump is P0 * envelope
    P0: float
    seed: int = 0
    psi0: np.ndarray | None = None  # warm start (ny, nx) complex
    res0: tuple | None = None       # warm start reservoir fields (ny, nx)
    label: dict = field(default_factory=dict)


@dataclass
class SweepResult:
    stop_times: np.ndarray          # ps; NaN for members that did not converge by max_time
    wall_s: float
    steps_run: int
    checks: int
    point_steps: float              # sum over members of grid points x steps actually stepped
    final_psi: list
    final_res: list
    relative_density: np.ndarray    # last relative density change per member
    wall_to_last_s: float = float("nan")  # wall time from launch to the last convergence
    wall_at_stop: np.ndarray | None = None  # s from launch to the check that found each member converged


class PackedSweep:
    def __init__(self, members: list[SweepMember], *, reservoir="single", precision="single", dt=0.05,
                 check_every=1.0, rtol=1e-6, min_time=0.0, max_time=1000.0, tile_h="auto", form="auto",
                 initial_amplitude=1e-3, retire=True, graphs=True, progress=None, progress_every_s=30.0):
        """``form`` is "step" (whole-step kernel), "stage" (four stage kernels) or "auto";
        ``tile_h`` a tile height or "auto", which asks the autotuner (cached per size mix).
        An explicit tile height with ``form="auto"`` takes the whole-step kernel where that
        height allows it (16, 32, 64)."""
        init()
        self.members = members
        self.reservoir = reservoir
        self.precision = precision
        self.dt = float(dt)
        self.steps_per_check = int(round(check_every / dt))
        if self.steps_per_check < 1 or abs(self.steps_per_check * dt - check_every) > 1e-9 * check_every:
            raise ValueError("check_every must be a multiple of dt")
        self.check_every = check_every
        self.rtol = rtol
        self.min_time = min_time
        self.max_checks = int(round(max_time / check_every))
        self.retire = retire
        self.graphs = graphs
        self.amplitude = initial_amplitude
        self.progress = progress
        self.progress_every_s = progress_every_s
        self.stream = Stream()
        specs = []
        off = 0
        for m in members:
            specs.append(MemberSpec(nx=m.nx, ny=m.ny, dx=m.lx / m.nx, dy=m.ly / m.ny, physics=m.physics,
                                    reservoir=reservoir, pump_off=off, seed=m.seed, dt=self.dt))
            off += m.nx * m.ny
        if tile_h == "auto":
            from .autotune import tune_rk4
            choice = tune_rk4(specs, precision, reservoir)
            if form in ("auto", choice.form):
                form, tile_h = choice.form, choice.tile_h
            else:
                tile_h = 32 if form == "step" else 16
        elif form == "auto":
            form = "step" if tile_h in (16, 32, 64) else "stage"
        self.form, self.tile_h = form, int(tile_h)
        self.table = MemberTable(specs, self.tile_h)
        r = real_dtype(precision)
        env = np.concatenate([np.asarray(m.envelope, dtype=np.float64).reshape(-1) for m in members])
        self.pump = DeviceArray.from_numpy(env.astype(r))
        self.engine = Rk4Engine(self.table, self.pump, precision, reservoir, self.stream,
                                whole_step=(form == "step"))
        self.engine.set_dt(self.dt)
        coef = np.tile(np.asarray([m.P0 for m in members], dtype=np.float64), (4, 1)).astype(r)
        self.coef = DeviceArray.from_numpy(coef)
        self._load_initial()
        n = self.table.total_points
        self.old = DeviceArray.zeros((n,), r)
        self.partials = DeviceArray.zeros((3 * self.table.num_tiles,), np.float64)
        self.sums = DeviceArray.zeros((3 * len(members),), np.float64)

    def _load_initial(self):
        c, r = complex_dtype(self.precision), real_dtype(self.precision)
        psi = []
        res = [[] for _ in range(self.engine.nres)]
        for m in self.members:
            if m.psi0 is not None:
                psi.append(np.asarray(m.psi0, dtype=c).reshape(-1))
            else:
                psi.append(make_initial_psi((m.ny, m.nx), self.amplitude, seed=m.seed,
                                            precision=self.precision).reshape(-1))
            for f in range(self.engine.nres):
                if m.res0 is not None:
                    res[f].append(np.maximum(np.asarray(m.res0[f], dtype=r).reshape(-1), 0))
                else:
                    res[f].append(np.zeros(m.nx * m.ny, dtype=r))
        self.engine.psi[0].upload(np.concatenate(psi))
        for f in range(self.engine.nres):
            self.engine.res[f].upload(np.concatenate(res[f]))

    def _density_check(self):
        e = self.engine
        prec = 0 if self.precision == "single" else 1
        N.call("pgpe_density_delta", e.psi[0].ptr, self.old.ptr, prec, 1, self.table.table_ptr, e.tile_map.ptr,
               e.num_tiles, self.table.members, self.partials.ptr, self.sums.ptr, 1, self.stream.handle)
        return self.sums.get(self.stream).reshape(-1, 3)

    def run(self) -> SweepResult:
        M = len(self.members)
        cells = np.asarray([m.lx / m.nx * m.ly / m.ny for m in self.members])
        points = np.asarray([m.nx * m.ny for m in self.members], dtype=np.float64)
        active = np.ones(M, dtype=bool)
        stop = np.full(M, np.nan)
        wall_at_stop = np.full(M, np.nan)
        rel_d_last = np.full(M, np.inf)
        prev_norm = None
        point_steps = 0.0
        self.stream.synchronize()
        t0 = time.perf_counter()
        step = 0
        checks = 0
        last_conv = float("nan")
        next_report = t0 + self.progress_every_s
        for c in range(self.max_checks):
            self.engine.run_steps(self.steps_per_check, self.coef, True, False, step, self.graphs)
            point_steps += float(points[active].sum()) * self.steps_per_check
            step += self.steps_per_check
            sums = self._density_check()
            checks += 1
            t = step * self.dt
            norms = sums[:, 0] * cells
            if prev_norm is not None:
                rel_d = np.sqrt(sums[:, 1]) / np.maximum(np.sqrt(sums[:, 2]), 1e-30)
                rel_n = np.abs(norms - prev_norm) / np.maximum(prev_norm, 1e-30)
                reached = (norms < 1e-30) | ((rel_n < self.rtol) & (rel_d < self.rtol))
                rel_d_last = np.where(active, rel_d, rel_d_last)
                if t >= self.min_time:
                    newly = reached & active
                    stop[newly] = t
                    if newly.any():
                        last_conv = time.perf_counter() - t0
                        wall_at_stop[newly] = last_conv
                    if newly.any() and self.retire:
                        active &= ~newly
                        self.engine.set_active_tiles(self.table.tiles_of(np.flatnonzero(active)))
                    elif newly.any():
                        active &= ~newly
            prev_norm = norms
            if self.progress is not None and time.perf_counter() > next_report:
                next_report = time.perf_counter() + self.progress_every_s
                self.progress(f"sweep t={t:.0f} ps: {int(np.isfinite(stop).sum())}/{M} converged, "
                              f"{int(active.sum())} running, {time.perf_counter() - t0:.0f} s")
            if not active.any():
                break
        self.stream.synchronize()
        wall = time.perf_counter() - t0
        psi_all = self.engine.psi[0].get()
        res_all = [r.get() for r in self.engine.res]
        final_psi, final_res = [], []
        for i, m in enumerate(self.members):
            o = int(self.table.offsets[i])
            n = m.nx * m.ny
            final_psi.append(psi_all[o:o + n].reshape(m.ny, m.nx))
            final_res.append(tuple(r[o:o + n].reshape(m.ny, m.nx) for r in res_all))
        return SweepResult(stop, wall, step, checks, point_steps, final_psi, final_res, rel_d_last, last_conv,
                           wall_at_stop)
