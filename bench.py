#!/usr/bin/env python
"""Benchmark: steady-state sweep members/s on one MI355X, plus the stencil roofline.

Workload (one "step" = one pass over the whole batch):
  256 held-out sweep members, 256 x 256 periodic grids (lx = ly = 128 um),
  single reservoir, complex64, rk4-fdm-hip with dt = 0.05 ps, random pumps
  (1-4 Gaussians / ring / square lattice, seeded) with peak P0 = U[0.8, 4] x the
  uniform-pump threshold.  Each member runs from a noise-seeded cold start until
  the reference's steady-state criterion holds (relative L2 density change and
  relative norm change < 1e-6 between checks 1 ps apart) or 4000 ps elapse.
  Converged members are retired from the launches.  value = converged members
  per second of wall time, launch to last member done.

Roofline leg: the middle RK4 stage kernel (FP32, single reservoir) on one
8192 x 8192 member (537 MB per complex field, far beyond the 256 MB Infinity
Cache), timed with HIP events on the solver stream.  Algorithmic bytes per
point for that stage: psi0, k_prev (2 x 8 B), n0, kn_prev, P (3 x 4 B) in;
k (8 B), kn (4 B) out = 40 B.

cpu_baseline: the NumPy oracle's rk4-fdm step (complex64, one 256 x 256
member) timed on one host core for a bounded number of steps; its point-step
rate converts the GPU run's point-steps into the CPU time for the same sweep.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SWEEP = dict(members=256, n=256, L=128.0, dt=0.05, check_every=1.0, rtol=1e-6, max_time=4000.0, seed=777)


def held_out_members(count, seed):
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.pumps import random_pump, threshold
    from polarism_amd.sweep import SweepMember
    p = PhysicsConstants()
    rng = np.random.default_rng(seed)
    pth = threshold(p, "single")
    n, L = SWEEP["n"], SWEEP["L"]
    out = []
    for i in range(count):
        env, desc = random_pump(rng, n, n, L, L)
        ratio = float(rng.uniform(0.8, 4.0))
        out.append(SweepMember(n, n, L, L, p, env, ratio * pth, seed=seed * 100003 + i,
                               label={"ratio": ratio, "kind": desc["kind"]}))
    return out


def run_sweep(members, tile_h=8):
    from polarism_amd.sweep import PackedSweep
    sw = PackedSweep(members, reservoir="single", precision="single", dt=SWEEP["dt"],
                     check_every=SWEEP["check_every"], rtol=SWEEP["rtol"], max_time=SWEEP["max_time"],
                     tile_h=tile_h)
    return sw.run()


def roofline_leg(n=8192, tile_h=16, reps=20):
    """Average duration of the FP32 middle-stage kernel on an n x n member."""
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.device import DeviceArray, Event
    from polarism_amd.engine import Rk4Engine
    from polarism_amd.layout import MemberSpec, MemberTable
    p = PhysicsConstants()
    tab = MemberTable([MemberSpec(n, n, 0.5, 0.5, p, reservoir="single", dt=0.05)], tile_h)
    pump = DeviceArray.from_numpy(np.full(n * n, 0.01, dtype=np.float32))
    e = Rk4Engine(tab, pump, "single", "single")
    e.set_dt(0.05)
    rng = np.random.default_rng(0)
    e.psi[0].upload((1e-3 * (rng.standard_normal(n * n) + 1j * rng.standard_normal(n * n))).astype(np.complex64))
    for buf in e.k + [e.psi[1]] + [x for row in e.kres for x in row]:
        buf.fill_zero(e.stream)
    coef = DeviceArray.from_numpy(np.ones(4, dtype=np.float32))
    ptrs = e.coef_ptrs(coef, 0)
    e.launch_step(e.psi[0], e.psi[1], ptrs, False, 0)
    e.stream.synchronize()
    start, stop = Event(), Event()
    total = 0.0
    for _ in range(reps):
        start.record(e.stream)
        e.launch_step(e.psi[0], e.psi[1], ptrs, False, 0, stages=(1,))
        stop.record(e.stream)
        total += start.elapsed_ms(stop)
    avg_ms = total / reps
    bytes_per_launch = 40.0 * n * n
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"rk4_stage_kernel<float, SINGLE, {tile_h}, middle>", "grid": f"{n}x{n}",
            "avg_us": round(avg_ms * 1e3, 2), "algorithmic_bytes_per_launch": int(bytes_per_launch)}


def cpu_baseline(point_steps, converged, steps=300):
    """Oracle rk4-fdm on the host (one core, complex64 256^2) scaled to the sweep's point-steps."""
    from oracle import gpe_oracle as O
    g = O.Grid(SWEEP["n"], SWEEP["n"], SWEEP["L"], SWEEP["L"])
    p = O.Physics()
    env = O.gaussian_profile(g, 12.0).astype(np.float32)
    pump_at = lambda t: env * np.float32(0.03)
    psi = O.initial_psi(g.shape, 1e-3, 0, "single").reshape(1, 1, *g.shape)
    res = (np.zeros((1, 1, *g.shape), np.float32),)
    psi, res = O.rk4_fdm_step("single", p, g, psi, res, pump_at, 0.0, SWEEP["dt"])
    t0 = time.perf_counter()
    for s in range(steps):
        psi, res = O.rk4_fdm_step("single", p, g, psi, res, pump_at, 0.0, SWEEP["dt"])
    dt = time.perf_counter() - t0
    rate = steps * g.nx * g.ny / dt
    cpu_time = point_steps / rate
    return {"value": round(converged / cpu_time, 6), "unit": "members/s", "cores": 1, "kind": "port",
            "sample": f"{steps} oracle rk4-fdm steps, complex64 256x256, {dt:.1f} s on one core "
                      f"({rate / 1e6:.2f} Mpt-step/s), scaled to the sweep's {point_steps:.3e} point-steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--members", type=int, default=SWEEP["members"])
    ap.add_argument("--legs", default="sweep,roofline,cpu")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (the sweep would shard "
              "members across GPUs; no multi-GPU run is part of it)", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    from polarism_amd.device import init
    init()
    result = {"metric": "steady-state sweep members/s (warm vs plain) at <=1e-3 rel L2; % HBM roofline",
              "unit": "members/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
              "data": "synthetic (seeded random pumps, noise-seeded initial fields)"}
    members = held_out_members(args.members, SWEEP["seed"])
    if "sweep" in legs:
        warm_members = held_out_members(min(16, args.members), SWEEP["seed"] + 1)
        for _ in range(args.warmup):
            run_sweep(warm_members)
        t0 = time.perf_counter()
        runs = [run_sweep(members) for _ in range(args.steps)]
        wall = time.perf_counter() - t0
        res = runs[-1]
        converged = int(np.isfinite(res.stop_times).sum())
        value = converged * args.steps / wall
        result.update(value=round(value, 4), ms_per_step=round(wall / args.steps * 1e3, 1))
        eff_bytes = res.point_steps * 172.0
        result["sweep"] = {
            "plain": {"members_per_s": round(value, 4), "converged": converged, "members": len(members),
                      "wall_s": round(wall / args.steps, 3), "steps_run": res.steps_run,
                      "median_stop_ps": float(np.nanmedian(res.stop_times)) if converged else None,
                      "point_steps": res.point_steps,
                      "effective_GBps": round(eff_bytes / (wall / args.steps) / 1e9, 1)},
            "warm": None}
    if "roofline" in legs:
        result["roofline"] = roofline_leg()
    result["config"] = {"workload": "steady-state pump sweep, 256 members of 256x256, rk4-fdm-hip complex64, "
                                    "single reservoir, dt 0.05 ps, cold start",
                        "members": len(members), "grid": "256x256", "lx_um": SWEEP["L"], "dt_ps": SWEEP["dt"],
                        "steady_rtol": SWEEP["rtol"], "check_every_ps": SWEEP["check_every"],
                        "max_time_ps": SWEEP["max_time"], "parallelism": "1 GPU, members packed in one launch"}
    if "cpu" in legs and "sweep" in result:
        s = result["sweep"]["plain"]
        result["cpu_baseline"] = cpu_baseline(s["point_steps"], max(s["converged"], 1))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
