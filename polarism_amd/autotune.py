"""Launch autotuner for the RK4 kernels, persisted per (grid set, method).

For a batch of members it times every launch form the device supports --
the four fused stage launches (1-point halo, tile heights 8 / 16 / 32) and
the whole-step launch (4-point halo recomputed on chip, LDS tile heights 16 / 32
and 64 in single precision, or register row strips 16 / 24 and 32 / 40 in
single precision high) -- on the batch's own member table, and keeps the
fastest.  Every form gives the same result (bit for bit on periodic grids,
to rounding on closed intervals; the kernels share their arithmetic), so the
choice only changes speed.  Choices are stored in a
JSON file keyed by device, method, precision, reservoir, components and the
multiset of grid sizes; later runs on the same key skip the timing.

The reference tunes only the batch size of a sweep (sweep.py:23-48) and picks
CUDA tiles by a shared-memory fit (rk4_cuda_solver.py:1060-1095).
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np


@dataclass(frozen=True)
class LaunchChoice:
    form: str      # "step" (whole-step kernel on LDS tiles), "strip" (whole-step kernel on
    tile_h: int    # register row strips) or "stage" (four stage kernels)

    @property
    def whole_step(self) -> bool:
        return self.form in ("step", "strip")

    @property
    def table_h(self) -> int:
        """Tile height of the member table (row strips keep the table at 16 for the reductions)."""
        return 16 if self.form == "strip" else self.tile_h

    @property
    def strip_h(self) -> int | None:
        return self.tile_h if self.form == "strip" else None


def candidates(precision: str, components: int) -> list[LaunchChoice]:
    out = [LaunchChoice("stage", h) for h in (8, 16, 32)]
    if components == 1:
        out += [LaunchChoice("step", h) for h in (16, 32)]
        out += [LaunchChoice("strip", h) for h in (8, 16, 24)]
        if precision == "single":
            out += [LaunchChoice("step", 64), LaunchChoice("strip", 32), LaunchChoice("strip", 40)]
    return out


def cache_path() -> Path:
    env = os.environ.get("PGPE_AUTOTUNE_CACHE")
    if env:
        return Path(env)
    return Path(os.path.expanduser("~")) / ".cache" / "polarism_amd" / "autotune.json"


def tune_key(method: str, device: str, precision: str, reservoir: str, components: int, sizes) -> str:
    counts: dict = {}
    for nx, ny in sizes:
        counts[(int(nx), int(ny))] = counts.get((int(nx), int(ny)), 0) + 1
    grids = ",".join(f"{nx}x{ny}*{c}" for (nx, ny), c in sorted(counts.items()))
    digest = hashlib.sha1(grids.encode()).hexdigest()[:16]
    return f"{method}|{device}|{precision}|{reservoir}|c{components}|{digest}"


class Autotuner:
    def __init__(self, path: Path | None = None):
        self.path = Path(path) if path is not None else cache_path()
        self._table = self._load()

    def _load(self) -> dict:
        try:
            with open(self.path) as f:
                data = json.load(f)
            return data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            return {}

    def _store(self) -> None:
        self.path.parent.mkdir(parents=True, exist_ok=True)
        tmp = self.path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(self._table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def lookup(self, key: str) -> LaunchChoice | None:
        v = self._table.get(key)
        if not v:
            return None
        return LaunchChoice(v["form"], int(v["tile_h"]))

    def choose(self, key: str, options: list[LaunchChoice], timer) -> LaunchChoice:
        """Cached choice for ``key``, else the option with the smallest ``timer(option)`` (seconds)."""
        hit = self.lookup(key)
        if hit is not None and hit in options:
            return hit
        times = {}
        for opt in options:
            times[opt] = float(timer(opt))
        best = min(options, key=lambda o: times[o])
        self._table[key] = {"form": best.form, "tile_h": best.tile_h,
                            "times_us": {f"{o.form}/{o.tile_h}": round(t * 1e6, 2) for o, t in times.items()},
                            "tuned_at": time.strftime("%Y-%m-%dT%H:%M:%S")}
        self._store()
        return best


def sample_members(specs, budget: int):
    """All of ``specs`` if their points fit ``budget``, else members at evenly spaced positions
    (keeping the size mix), as many as fit ``budget`` (at least one)."""
    total = sum(s.nx * s.ny for s in specs)
    if total <= budget:
        return list(specs)
    count = max(1, int(len(specs) * budget / total))
    while count > 1:
        idx = np.linspace(0, len(specs) - 1, count).round().astype(int)
        if sum(specs[i].nx * specs[i].ny for i in idx) <= budget:
            break
        count -= 1
    idx = np.linspace(0, len(specs) - 1, count).round().astype(int)
    return [specs[i] for i in idx]


def tune_rk4(specs, precision: str, reservoir: str, *, steps: int = 20, tuner: Autotuner | None = None,
             sample_points: int = 1 << 25, options: list[LaunchChoice] | None = None) -> LaunchChoice:
    """Fastest RK4 launch form and tile height for the members ``specs`` (MemberSpec list).

    Times ``steps`` (even) graph-captured steps per candidate, twice, keeping the faster, on the
    whole member table when it holds at most ``sample_points`` points, else on members drawn
    evenly from ``specs`` up to that many points.  The table has to be near full size: on a
    fraction of it the persistent strip kernels run a few strips per workgroup and short strips
    time like tall ones, though on the full table the tall ones are faster.  ``options`` limits
    the forms (default: every candidate of the precision and component count); a cached choice
    outside them is timed again.
    """
    from .device import DeviceArray, Event, device_name, init
    from .engine import Rk4Engine, real_dtype
    from .layout import MemberTable
    init()
    components = specs[0].components
    sizes = [(s.nx, s.ny) for s in specs]
    key = tune_key("rk4-fdm-hip", device_name(), precision, reservoir, components, sizes)
    tuner = tuner or Autotuner()
    opts = list(options) if options is not None else candidates(precision, components)
    hit = tuner.lookup(key)
    if hit is not None and hit in opts:
        return hit
    sub = sample_members(specs, sample_points)
    off = 0
    sub_specs = []
    for s in sub:
        sub_specs.append(type(s)(**{**s.__dict__, "pump_off": off}))
        off += s.nx * s.ny
    r = real_dtype(precision)
    rng = np.random.default_rng(0)
    pump = DeviceArray.from_numpy(rng.uniform(0, 1, off).astype(r))
    coef = DeviceArray.from_numpy(np.full(4 * len(sub_specs), 0.03, dtype=r))
    n_field = off * components
    psi = (1e-2 * (rng.standard_normal(n_field) + 1j * rng.standard_normal(n_field)))

    def timer(opt: LaunchChoice) -> float:
        tab = MemberTable(sub_specs, opt.table_h)
        e = Rk4Engine(tab, pump, precision, reservoir, whole_step=opt.whole_step, strip_h=opt.strip_h)
        e.set_dt(sub_specs[0].dt)
        e.psi[0].upload(psi.astype(e.cdtype))
        e.run_steps(steps, coef, True, False, 0)      # captures the graph and warms up
        best = float("inf")
        for _ in range(2):
            start, stop = Event(), Event()
            start.record(e.stream)
            e.run_steps(steps, coef, True, False, 0)
            stop.record(e.stream)
            e.stream.synchronize()
            best = min(best, start.elapsed_ms(stop) * 1e-3 / steps)
        return best

    return tuner.choose(key, opts, timer)
