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
    boundary: str = "periodic"      # or "closed-interval" (dx = lx / (nx - 1))
    laplacian: str = "five-point"   # or "isotropic-9pt"
    noise_strength: float = 0.0     # truncated-Wigner strength (0: no noise)
    initial_amplitude: float | None = None  # cold-start RMS; None: the sweep's default

    @property
    def spacing(self) -> tuple[float, float]:
        if self.boundary == "closed-interval":
            return self.lx / (self.nx - 1), self.ly / (self.ny - 1)
        return self.lx / self.nx, self.ly / self.ny


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
        """``form`` is "step" (whole-step kernel on LDS tiles), "strip" (whole-step kernel on
        register row strips), "stage" (four stage kernels) or "auto"; ``tile_h`` a tile (strip)
        height or "auto", which asks the autotuner (cached per size mix).
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
            dx, dy = m.spacing
            specs.append(MemberSpec(nx=m.nx, ny=m.ny, dx=dx, dy=dy, physics=m.physics, reservoir=reservoir,
                                    pump_off=off, seed=m.seed, dt=self.dt, noise_strength=m.noise_strength,
                                    laplacian=m.laplacian, boundary=m.boundary))
            off += m.nx * m.ny
        self.noise = any(m.noise_strength > 0 for m in members)
        if tile_h == "auto":
            from .autotune import tune_rk4
            choice = tune_rk4(specs, precision, reservoir)
            if form in ("auto", choice.form):
                form, tile_h = choice.form, choice.tile_h
            else:
                tile_h = {"step": 32, "strip": 24}.get(form, 16)
        elif form == "auto":
            form = "step" if tile_h in (16, 32, 64) else "stage"
        if form not in ("step", "strip", "stage"):
            raise ValueError(f"unknown launch form {form!r}")
        self.form, self.tile_h = form, int(tile_h)
        self.table = MemberTable(specs, 16 if form == "strip" else self.tile_h)
        r = real_dtype(precision)
        env = np.concatenate([np.asarray(m.envelope, dtype=np.float64).reshape(-1) for m in members])
        self.pump = DeviceArray.from_numpy(env.astype(r))
        self.engine = Rk4Engine(self.table, self.pump, precision, reservoir, self.stream,
                                whole_step=form in ("step", "strip"),
                                strip_h=self.tile_h if form == "strip" else None)
        self.engine.set_dt(self.dt)
        coef = np.tile(np.asarray([m.P0 for m in members], dtype=np.float64), (4, 1)).astype(r)
        self.coef = DeviceArray.from_numpy(coef)
        self._load_initial()
        n = self.table.total_points
        self.old = DeviceArray.zeros((n,), r)
        self.partials = DeviceArray.zeros((3 * self.table.num_tiles,), np.float64)
        # the convergence check reduces over the member table's tiles of the running members
        # (its own map: a strip launch map is cut differently)
        self.check_map = DeviceArray.from_numpy(self.table.tile_map_host.reshape(-1))
        self.check_tiles = self.table.num_tiles
        self.sums = DeviceArray.zeros((3 * len(members),), np.float64)

    def _load_initial(self):
        c, r = complex_dtype(self.precision), real_dtype(self.precision)
        psi = []
        res = [[] for _ in range(self.engine.nres)]
        for m in self.members:
            if m.psi0 is not None:
                psi.append(np.asarray(m.psi0, dtype=c).reshape(-1))
            else:
                amp = self.amplitude if m.initial_amplitude is None else m.initial_amplitude
                psi.append(make_initial_psi((m.ny, m.nx), amp, seed=m.seed, precision=self.precision).reshape(-1))
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
        N.call("pgpe_density_delta", e.psi[0].ptr, self.old.ptr, prec, 1, self.table.table_ptr, self.check_map.ptr,
               self.check_tiles, self.table.members, self.partials.ptr, self.sums.ptr, 1, self.stream.handle)
        return self.sums.get(self.stream).reshape(-1, 3)

    def run(self) -> SweepResult:
        M = len(self.members)
        cells = np.asarray([m.spacing[0] * m.spacing[1] for m in self.members])
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
            self.engine.run_steps(self.steps_per_check, self.coef, True, self.noise, step, self.graphs)
            point_steps += float(points[active].sum()) * self.steps_per_check
            step += self.steps_per_check
            sums = self._density_check()
            checks += 1
            t = step * self.dt
            norms = sums[:, 0] * cells
            if prev_norm is not None and self.rtol is not None:
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
                        running = np.flatnonzero(active)
                        self.engine.set_active_members(running)
                        pairs = self.table.tiles_of(running).reshape(-1)
                        if pairs.size:
                            self.check_map.view_range(0, (pairs.size,)).upload(pairs, self.stream)
                        self.check_tiles = pairs.size // 2
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


# ---------------------------------------------------------------------------- Sweep
# The reference's parameter-sweep entry point, Sweep(cfg, parameters, mode).run()
# (sweep.py:98-324 with the override paths of sweep_config.py:31-124).  Members are the
# Cartesian product (or zip) of the parameter sequences, in order.  Where the reference runs
# one batch per group of members that share the "group" settings (grid, solver, reservoir, ...),
# members that rk4-fdm-hip can step together -- any grid sizes, the same reservoir model,
# precision and time step, time-independent pumps -- go into one packed launch here; the rest
# run as controller batches, one per group, as in the reference.

_GROUP_SECTIONS = frozenset({"grid", "solver", "reservoir", "boundary_condition", "compute_engine", "result",
                             "checkpoint", "stop", "sweep", "terms"})
_GROUP_FIELDS = {"physics": {"dispersion", "spin"}, "noise": {"enabled", "stochastic"},
                 "initial_condition": {"type", "path"}, "laser": {"laser_type", "mode"},
                 "potential": {"potential_type"}, "drive": {"enabled"}}
# group keys that the packed launch still runs side by side (members keep their own grids)
_PACKABLE_GROUP_KEYS = frozenset({"grid.nx", "grid.ny", "grid.lx", "grid.ly"})


def override_kind(path: str) -> str:
    """'batch' if members differing in ``path`` can share one controller batch, 'group' if they
    need separate ones; ValueError for a path that names no configuration key."""
    from .config import Config
    section, _, key = path.partition(".")
    if not key or section not in Config._declared:
        raise ValueError(f"Unknown sweep path {path!r}; use paths such as physics.g_C, laser.P0 or grid.nx")
    if section == "terms":
        return "group"
    sec = getattr(Config(), section)
    head, _, sub = key.partition(".")
    if head not in type(sec)._declared:
        raise ValueError(f"Unknown sweep path {path!r}")
    if sub and not isinstance(getattr(sec, head), dict):
        raise ValueError(f"Unknown sweep path {path!r}; {section}.{head} is not a mapping")
    if section in _GROUP_SECTIONS:
        return "group"
    if sub:   # profile / envelope / extra entries: the type selects a different function
        return "group" if sub == "type" else "batch"
    return "group" if head in _GROUP_FIELDS.get(section, ()) else "batch"


def _manifest_value(value):
    if hasattr(value, "item"):
        return value.item()
    raise TypeError(f"Sweep value {value!r} cannot be encoded as JSON")


def _config_signature(cfg) -> dict:
    """The settings that determine the stored frames (the reference's manifest field)."""
    import json
    from dataclasses import asdict
    data = asdict(cfg)
    keep = ("grid", "solver", "reservoir", "physics", "terms", "initial_condition", "laser", "potential", "drive",
            "boundary_condition", "stop", "compute_engine")
    return json.loads(json.dumps({k: data[k] for k in keep}, default=_manifest_value))


def _packable(cfg) -> bool:
    """Can PackedSweep step this member: rk4-fdm-hip on a scalar field, fixed steps, no
    potential or drive, one laser whose coefficient does not change with time."""
    from .pump import LASERS
    laser = cfg.laser
    if cfg.solver.method != "rk4-fdm-hip" or cfg.solver.adaptive or cfg.physics.spin != "scalar":
        return False
    if cfg.potential.potential_type != "zero" or cfg.drive.enabled or cfg.terms:
        return False
    if laser.mode != "single" or laser.delay != 0.0 or laser.laser_type not in LASERS:
        return False
    if not getattr(LASERS[laser.laser_type], "time_independent", False):
        return False
    return cfg.noise.stochastic in ("none", "wigner") and cfg.result.save_every is None


class Sweep:
    """Run the members of a parameter sweep (sweep.py:98-324).

    ``parameters`` maps dotted configuration paths (``physics.g_C``, ``laser.P0``,
    ``grid.nx`` ...) to sequences; ``mode`` is "product" or "zip".  ``run()`` returns one
    dict per member, in member order, with the final ``psi`` (public shape), ``reservoir``
    fields and, for the packed launch, the steady-state ``stop_time`` (NaN if not reached).
    With ``result.save_results`` or a checkpoint path a run directory is created holding
    ``sweep.json`` (the reference's manifest: format_version, mode, parameter_keys, parameters,
    members -> (group, batch, position), batches, batch_sizes, config_signature, seed,
    member_seeds) and each batch's final fields in ``batch_NNN/final.npz``."""

    def __init__(self, cfg, parameters, mode="product"):
        from itertools import product
        if mode not in ("product", "zip"):
            raise ValueError("mode must be 'product' or 'zip'")
        if not isinstance(parameters, dict):
            raise ValueError("parameters must map dotted paths to sequences")
        keys = list(parameters)
        kinds = {k: override_kind(k) for k in keys}
        columns = [list(parameters[k]) for k in keys]
        if any(len(c) == 0 for c in columns):
            raise ValueError("sweep parameter sequences must be nonempty")
        if mode == "zip" and len({len(c) for c in columns}) > 1:
            raise ValueError("zip sweep parameter sequences must have equal lengths")
        rows = product(*columns) if mode == "product" else zip(*columns)
        self.cfg, self.mode, self.kinds = cfg, mode, kinds
        self.members = [dict(zip(keys, row)) for row in rows]
        if cfg.result.view_member >= len(self.members):
            raise ValueError("result.view_member is outside the sweep")
        self.batches = self._plan()

    # -- planning ---------------------------------------------------------------------
    def _member_cfg(self, overrides):
        from .controller import apply_overrides
        from .validation import validate_config
        out = apply_overrides(self.cfg, overrides)
        validate_config(out)
        return out

    def _plan(self) -> list[dict]:
        """Batches in run order: {"group": g, "indices": [...], "packed": bool}.  Packed batches
        gather every packable member with the same reservoir model, precision, time step, stop
        rule and noise kind, whatever their grids; the other members form one batch per group
        of equal group settings, as in the reference."""
        import json
        order, batches = {}, []
        for index, overrides in enumerate(self.members):
            cfg = self._member_cfg(overrides)
            if _packable(cfg):
                s = cfg.solver
                key = ("packed", cfg.reservoir.reservoir_type, s.precision, s.dt, s.laplacian, cfg.grid.grid_type,
                       cfg.noise.stochastic, json.dumps(_config_signature(cfg)["stop"]), s.total_time)
            else:
                key = ("group", json.dumps({k: v for k, v in overrides.items() if self.kinds[k] == "group"},
                                           sort_keys=True, default=str))
            if key not in order:
                order[key] = len(batches)
                batches.append({"group": len(batches), "indices": [], "packed": key[0] == "packed"})
            batches[order[key]]["indices"].append(index)
        for b in batches:
            b["size"] = len(b["indices"])
        return batches

    def manifest(self, seed, member_seeds) -> dict:
        from . import __version__
        mapping = [None] * len(self.members)
        for bi, b in enumerate(self.batches):
            for pos, index in enumerate(b["indices"]):
                mapping[index] = {"group": b["group"], "batch": bi, "position": pos}
        return {"format_version": 1, "polarism_version": __version__, "mode": self.mode,
                "parameter_keys": list(self.members[0]), "parameters": self.members, "members": mapping,
                "batches": [{"group": b["group"], "indices": b["indices"], "size": b["size"], "packed": b["packed"]}
                            for b in self.batches],
                "batch_sizes": [b["size"] for b in self.batches], "config_signature": _config_signature(self.cfg),
                "seed": seed, "member_seeds": member_seeds}

    # -- execution --------------------------------------------------------------------
    def _sweep_member(self, cfg, seed) -> SweepMember:
        from .grid import create_grid
        from .pump import create_lasers
        g = cfg.grid
        grid = create_grid(g)
        laser = create_lasers(cfg.laser, grid, "double")[0]
        noise_on = cfg.noise.enabled and cfg.noise.stochastic == "wigner"
        amp = cfg.noise.initial_amplitude if cfg.noise.enabled else 0.0
        return SweepMember(g.nx, g.ny, g.lx, g.ly, cfg.physics, laser.envelope, laser.power_coefficient(0.0),
                           seed=seed, boundary=g.grid_type, laplacian=cfg.solver.laplacian,
                           noise_strength=cfg.noise.stochastic_strength if noise_on else 0.0, initial_amplitude=amp)

    def _run_packed(self, batch, cfgs, seeds):
        members = [self._sweep_member(cfgs[i], seeds[i]) for i in batch["indices"]]
        c0 = cfgs[batch["indices"][0]]
        s, stop = c0.solver, c0.stop
        steady = stop.steady_state
        every = stop.check_every if steady else s.dt * max(1, min(1000, round(s.total_time / s.dt)))
        sw = PackedSweep(members, reservoir=c0.reservoir.reservoir_type, precision=s.precision, dt=s.dt,
                         check_every=every, rtol=stop.rtol if steady else None,
                         min_time=stop.min_time if steady else 0.0, max_time=s.total_time)
        res = sw.run()
        return [{"psi": res.final_psi[k], "reservoir": res.final_res[k], "stop_time": float(res.stop_times[k])}
                for k in range(len(members))]

    def _run_group(self, batch, cfgs):
        from .controller import SimulationController
        first = cfgs[batch["indices"][0]]
        base = self._member_cfg({k: v for k, v in self.members[batch["indices"][0]].items()
                                 if self.kinds[k] == "group"})
        local = [{k: v for k, v in self.members[i].items() if self.kinds[k] == "batch"} for i in batch["indices"]]
        base.result.save_results = False
        ctl = SimulationController(base, members=local if len(local) > 1 else (local or None))
        ctl.run()
        psi = ctl.state.psi
        fields = [r.get() for r in ctl.reservoir.get_state()]
        out = []
        for pos in range(len(local)):
            p_ = psi[pos] if ctl.state.batched else psi
            r_ = tuple((f[pos] if ctl.state.batched else f[0])[0] for f in fields)
            t = ctl.stop_times[pos] if first.stop.steady_state else np.nan
            out.append({"psi": np.array(p_), "reservoir": r_, "stop_time": float(t)})
        return out

    def run(self, resume=None):
        import json
        from pathlib import Path
        from .noise import draw_seed
        from .results import create_run_directory
        cfg = self.cfg
        seed = cfg.noise.seed if cfg.noise.seed is not None else draw_seed()
        seeds = [int(m.get("noise.seed", seed)) if m.get("noise.seed") is not None else int(seed)
                 for m in self.members]
        cfgs = []
        for m, sd in zip(self.members, seeds):
            c = self._member_cfg(m)
            c.noise.seed = sd
            cfgs.append(c)
        want_dir = cfg.result.save_results or cfg.checkpoint.path is not None or resume is not None
        run_dir = Path(resume) if resume is not None else (
            create_run_directory(cfg.result.output_directory) if want_dir else None)
        self.run_directory = run_dir
        manifest = self.manifest(int(seed), seeds)
        if run_dir is not None:
            path = run_dir / "sweep.json"
            if resume is not None:
                saved = json.loads(path.read_text())
                fresh = json.loads(json.dumps(manifest, default=_manifest_value))
                if saved["parameters"] != fresh["parameters"] or saved["mode"] != self.mode:
                    raise ValueError("sweep parameters or mode differ from the saved sweep")
                if saved["config_signature"] != fresh["config_signature"]:
                    raise ValueError("sweep simulation config differs from the saved sweep")
            else:
                path.write_text(json.dumps(manifest, indent=2, default=_manifest_value))
        results = [None] * len(self.members)
        for bi, batch in enumerate(self.batches):
            bdir = run_dir / f"batch_{bi:03d}" if run_dir is not None else None
            if bdir is not None and (bdir / "complete").exists():
                with np.load(bdir / "final.npz", allow_pickle=False) as z:
                    for pos, index in enumerate(batch["indices"]):
                        nres = int(z["reservoir_fields"])
                        results[index] = {"psi": z[f"psi_{pos}"], "stop_time": float(z["stop_time"][pos]),
                                          "reservoir": tuple(z[f"res_{pos}_{f}"] for f in range(nres))}
                continue
            out = self._run_packed(batch, cfgs, seeds) if batch["packed"] else self._run_group(batch, cfgs)
            for pos, index in enumerate(batch["indices"]):
                results[index] = out[pos]
            if bdir is not None:
                bdir.mkdir(parents=True, exist_ok=True)
                arrays = {"stop_time": np.asarray([o["stop_time"] for o in out]),
                          "reservoir_fields": np.asarray(len(out[0]["reservoir"]))}
                for pos, o in enumerate(out):
                    arrays[f"psi_{pos}"] = o["psi"]
                    for f, r in enumerate(o["reservoir"]):
                        arrays[f"res_{pos}_{f}"] = r
                np.savez(bdir / "final.npz", **arrays)
                (bdir / "complete").touch()
        return results
