#!/usr/bin/env python
"""Benchmark: steady-state sweep members/s on one MI355X, warm start vs plain.

Workload (one "step" = one pass over the whole held-out batch, plain and warm):
  256 held-out sweep members, 256 x 256 periodic grids (lx = ly = 128 um),
  single reservoir, complex64, rk4-fdm-hip with dt = 0.05 ps, seeded random
  pumps (1-4 Gaussians / ring / square lattice) with peak P0 = U[0.8, 4] x the
  uniform-pump threshold.  Each member runs until the reference's steady-state
  criterion holds (relative L2 density change and relative norm change both
  < 1e-6 between checks 1 ps apart) or 4000 ps elapse; converged members are
  retired from the launches.
  plain: noise-seeded cold start.
  warm : FNO surrogate prediction (psi, n_R); a member predicted to hold no
         condensate starts from psi = 0 (it meets the criterion at once, so it
         counts only where its cold start also decayed to the zero state).
  The gate: a member counts (for both forms) only if its plain run converged,
  its warm run converged, and their final densities agree within 1e-3
  relative L2; the plain run is the judge.
  members/s = gated members / wall time from launch to the last gated
  member's convergence (warm start includes the surrogate inference).
  value = warm members/s on the gated set; speedup = warm / plain on that set.
  A sample of gated members is rerun in complex128 and at half the time step
  (outside the timed steps) to report the FP32-vs-FP64 and dt sensitivity of
  the steady density.

Roofline leg: the headline's dominant kernel, the whole-step RK4 kernel (all
four stages of a step in one launch; FP32, single reservoir) in the form the
autotuner picks for the sweep (register row strips of 56 output columns, or
LDS halo tiles) on the sweep's 256-member table, timed with HIP events on the
engine's stream over back-to-back launches.  Its algorithmic bytes are psi0,
n0, P in and psi1, n1 out = 28 B per point-step.  Beside it: the middle RK4
stage kernel on one 8192 x 8192 member (537 MB per complex field, beyond the
256 MB Infinity Cache; psi0, k_prev, n0, kn_prev, P in, k, kn out = 40 B per
point) and both whole-step forms on that member.  HBM traffic counters come
from separate rocprofv3 --pmc runs (not part of this program), so ``traffic``
is null here.

Extra legs (reported beside the headline, not part of value):
  uneven:     256 members with grids drawn from {64..512}^2, double reservoir,
              complex64, dt 1e-3 ps, 2000 steps in 1000-step graph chunks
              (grid-point-stage updates/s).
  quickstart: the reference quickstart (rk4-fdm, 128^2, complex128, 2000
              steps of 1e-3 ps), steps/s.
  spinor:     spinor RK4 on one 1024^2 complex64 member: the whole-step
              kernel against the four stage launches (us per step).
  fft:        split-step FFT (CPU operator order, 4 transforms per step) and
              the K-outer order (2 per step), 512 x 512, single reservoir,
              Wigner noise seed 1234, dt 5e-3 ps, complex64 (steps/s and FFT
              GB/s with 2 x 8 B per point per transform).
  ffthbm:     batched forward 2-D FFT of 16 complex64 2048^2 planes.
  lawson:     spinor Lawson IFRK DP5, adaptive 1 ps ticks, 1024^2, double
              reservoir, complex64 (device and host trial controller), and
              the complex64 field against a complex128 run after 6 ps.
  fno:        the surrogate's BF16 MFMA GEMM (8192^3, B row-major and B with
              k contiguous) and its forward / training step at batch 16 on
              256^2 maps.

cpu_baseline: the NumPy oracle's rk4-fdm step (complex64, one 256 x 256
member) on one host core and on a pool of up to 16 processes (one member
each), each rate the median of three timings of a bounded number of steps;
the pooled rate converts the plain run's point-steps into CPU time.  The
quickstart, split-step and spinor Lawson configurations are timed with the
oracle on one core (median of three) and set beside their GPU legs.
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
WEIGHTS = os.path.join(ROOT, "polarism_amd", "surrogate", "weights", "fno_steady_256.npz")


def progress(msg):
    """One line on stderr when a step or a leg has finished (the JSON result alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_sweep(members, tile_h="auto", graphs=True):
    from polarism_amd.sweep import PackedSweep
    sw = PackedSweep(members, reservoir="single", precision="single", dt=SWEEP["dt"],
                     check_every=SWEEP["check_every"], rtol=SWEEP["rtol"], max_time=SWEEP["max_time"],
                     tile_h=tile_h, graphs=graphs, progress=lambda m: print(m, file=sys.stderr, flush=True))
    return sw.run()


def summarise(res, extra_s=0.0):
    conv = np.isfinite(res.stop_times)
    t_last = (res.wall_to_last_s if conv.any() else res.wall_s) + extra_s
    return {"members_per_s": round(int(conv.sum()) / t_last, 4) if conv.any() else 0.0,
            "converged": int(conv.sum()), "members": int(conv.size),
            "wall_to_last_s": round(t_last, 3), "wall_total_s": round(res.wall_s + extra_s, 3),
            "median_stop_ps": float(np.nanmedian(res.stop_times)) if conv.any() else None,
            "last_stop_ps": float(np.nanmax(res.stop_times)) if conv.any() else None,
            "steps_run": res.steps_run, "point_steps": res.point_steps,
            "point_steps_per_s": round(res.point_steps / res.wall_s, 1),
            "effective_GBps_172B_model": round(res.point_steps * 172.0 / res.wall_s / 1e9, 1)}


def _member_engine(n, tile_h, whole_step, reservoir="single", strip_h=None):
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.device import DeviceArray
    from polarism_amd.engine import Rk4Engine
    from polarism_amd.layout import MemberSpec, MemberTable
    p = PhysicsConstants()
    tab = MemberTable([MemberSpec(n, n, 0.5, 0.5, p, reservoir=reservoir, dt=0.05)], tile_h)
    pump = DeviceArray.from_numpy(np.full(n * n, 0.01, dtype=np.float32))
    e = Rk4Engine(tab, pump, "single", reservoir, whole_step=whole_step, strip_h=strip_h)
    e.set_dt(0.05)
    rng = np.random.default_rng(0)
    e.psi[0].upload((1e-3 * (rng.standard_normal(n * n) + 1j * rng.standard_normal(n * n))).astype(np.complex64))
    for buf in e.k + [e.psi[1]] + [x for row in e.kres for x in row]:
        buf.fill_zero(e.stream)
    coef = DeviceArray.from_numpy(np.ones(4, dtype=np.float32))
    return e, coef


def _time_launches(e, launch, reps):
    """Average duration of ``reps`` back-to-back launches, from HIP events on the launch stream.
    One launch is queued ahead of the start event, so the host's launch latency (the ctypes
    call) overlaps a running kernel instead of opening a gap inside the timed region."""
    from polarism_amd.device import Event
    launch()
    e.stream.synchronize()
    start, stop = Event(), Event()
    launch()
    start.record(e.stream)
    for _ in range(reps):
        launch()
    stop.record(e.stream)
    e.stream.synchronize()
    return start.elapsed_ms(stop) / reps


def sweep_step_kernel(reps=20):
    """Average duration of the whole-step kernel on the headline workload's member table (all 256
    members of 256^2 active, complex64, single reservoir, the autotuner's whole-step form: register
    row strips or LDS tiles, and its height), timed with
    HIP events on the engine's stream around ``reps`` direct launches.  Algorithmic bytes per
    point-step: psi0 (8), n0 (4), P (4) in, psi1 (8), n1 (4) out = 28 B."""
    from polarism_amd.autotune import tune_rk4
    from polarism_amd.device import DeviceArray
    from polarism_amd.engine import Rk4Engine
    from polarism_amd.layout import MemberSpec, MemberTable
    from polarism_amd.surrogate import warmstart as ws
    members = ws.make_members(SWEEP["members"], SWEEP["seed"])
    specs, off = [], 0
    for m in members:
        specs.append(MemberSpec(m.nx, m.ny, m.lx / m.nx, m.ly / m.ny, m.physics, pump_off=off, seed=m.seed,
                                dt=SWEEP["dt"]))
        off += m.nx * m.ny
    choice = tune_rk4(specs, "single", "single")
    if not choice.whole_step:
        choice = type(choice)("step", 32)
    tab = MemberTable(specs, choice.table_h)
    pump = DeviceArray.from_numpy(np.concatenate([m.envelope.reshape(-1) for m in members]).astype(np.float32))
    e = Rk4Engine(tab, pump, "single", "single", whole_step=True, strip_h=choice.strip_h)
    e.set_dt(SWEEP["dt"])
    rng = np.random.default_rng(0)
    e.psi[0].upload((0.1 * (rng.standard_normal(off) + 1j * rng.standard_normal(off))).astype(np.complex64))
    coef = DeviceArray.from_numpy(np.tile(np.asarray([m.P0 for m in members]), (4, 1)).astype(np.float32))
    ptrs = e.coef_ptrs(coef, 0)
    r0, r1 = e.res, e._res_alt
    ms = _time_launches(e, lambda: e.launch_step(e.psi[0], e.psi[1], ptrs, False, 0, res_in=r0, res_out=r1), reps)
    if choice.form == "strip":
        # strip height TH = NW R - 8 (wavefronts per workgroup x rows per wavefront, less the halo)
        nw, r = {8: (4, 4), 16: (4, 6), 24: (8, 4), 32: (8, 5), 40: (8, 6)}[choice.tile_h]
        name = f"rk4_strip_pf_kernel<1, {nw}, {r}, false, false, true>"
        shape = f"row strips of 56 x {choice.tile_h} output points"
    else:
        name = f"rk4_step_kernel<float, 1, {choice.tile_h}>"
        shape = f"LDS tiles of 64 x {choice.tile_h} output points"
    return ms, off, name, shape, e.capacity


def roofline_leg(n=8192, reps=20):
    """The roofline object is the headline's dominant kernel, the whole-step RK4 kernel on the
    sweep's 256-member table; the 8192^2 single-member figures (beyond the Infinity Cache) of
    the middle stage kernel and the whole-step kernel are reported beside it."""
    def stage_kernel_ms(th):
        e, coef = _member_engine(n, th, False)
        ptrs = e.coef_ptrs(coef, 0)
        e.launch_step(e.psi[0], e.psi[1], ptrs, False, 0)
        mid = _time_launches(e, lambda: e.launch_step(e.psi[0], e.psi[1], ptrs, False, 0, stages=(1,)), reps)
        whole = _time_launches(e, lambda: e.launch_step(e.psi[0], e.psi[1], ptrs, False, 0), reps)
        return mid, whole

    def step_kernel_ms(th, strip_h=None):
        e, coef = _member_engine(n, th, True, strip_h=strip_h)
        ptrs = e.coef_ptrs(coef, 0)
        r0, r1 = e.res, e._res_alt
        return _time_launches(e, lambda: e.launch_step(e.psi[0], e.psi[1], ptrs, False, 0, res_in=r0, res_out=r1),
                              reps)

    ms, points, name, shape, blocks = sweep_step_kernel(reps)
    per_launch = 28.0 * points
    achieved = per_launch / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "traffic_note": "no counters in this run; FETCH_SIZE / WRITE_SIZE of the same launches are in "
                            "profiles/rk4_strip_pf_kernel_sweep256.md",
            "kernel": name, "workload": f"256 members of 256x256 (sweep table), {shape}, {blocks} workgroups",
            "avg_us": round(ms * 1e3, 2), "algorithmic_bytes_per_launch": int(per_launch),
            "bytes_per_point_step": 28, "point_steps_per_s": round(points / (ms * 1e-3), 1),
            "effective_GBps_172B_per_stage_model": round(172.0 * points / (ms * 1e-3) / 1e9, 1)}
    stage = {th_: stage_kernel_ms(th_) for th_ in (16, 32)}
    best = min(stage, key=lambda t: stage[t][0])
    mid_ms, step_ms = stage[best]
    pts = float(n) * n
    roof["stage_kernel_8192"] = {
        "kernel": f"rk4_stage_kernel<float, SINGLE, {best}, middle>", "avg_us": round(mid_ms * 1e3, 2),
        "algorithmic_bytes_per_launch": int(40.0 * pts),
        "frac_of_hbm_peak": round(40.0 * pts / (mid_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "avg_us_by_tile_h": {str(k): round(v[0] * 1e3, 2) for k, v in stage.items()},
        "four_stage_step_us": round(step_ms * 1e3, 2)}
    ws_ = {}
    for label, th_, sh in (("tile_h_32", 32, None), ("tile_h_64", 64, None), ("strip_h_24", 16, 24),
                           ("strip_h_32", 16, 32)):
        t = step_kernel_ms(th_, sh)
        ws_[label] = {"avg_us": round(t * 1e3, 2), "point_steps_per_s": round(pts / (t * 1e-3), 1),
                      "frac_of_hbm_peak_28B": round(28.0 * pts / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "effective_GBps_172B_per_stage_model": round(172.0 * pts / (t * 1e-3) / 1e9, 1)}
    roof["step_kernel_8192"] = ws_
    return roof


def copy_leg(nbytes=1 << 30, reps=10):
    """Device-to-device copy bandwidth (read + write bytes), the measured ceiling beside the 8 TB/s peak."""
    from polarism_amd.device import DeviceArray, Event, Stream
    s = Stream()
    a = DeviceArray.zeros((nbytes // 4,), np.float32)
    b = DeviceArray.zeros((nbytes // 4,), np.float32)
    b.copy_from(a, s)
    s.synchronize()
    start, stop = Event(), Event()
    start.record(s)
    for _ in range(reps):
        b.copy_from(a, s)
    stop.record(s)
    s.synchronize()
    ms = start.elapsed_ms(stop) / reps
    return {"bytes_per_copy": nbytes, "avg_us": round(ms * 1e3, 2),
            "GBps_read_plus_write": round(2.0 * nbytes / (ms * 1e-3) / 1e9, 1)}


def uneven_leg(steps=2000, seed=4):
    """256 members of uneven grids, double reservoir, complex64, dt 1e-3 ps, 1000-step graph chunks."""
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.device import DeviceArray, Event
    from polarism_amd.engine import Rk4Engine
    from polarism_amd.layout import MemberSpec, MemberTable
    from polarism_amd.pumps import threshold
    from polarism_amd.autotune import tune_rk4
    p = PhysicsConstants()
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 96, 128, 192, 256, 384, 512], size=256)
    specs, envs, off = [], [], 0
    pth = threshold(p, "double")
    for i, nn in enumerate(sizes):
        nn = int(nn)
        specs.append(MemberSpec(nn, nn, 0.5, 0.5, p, reservoir="double", pump_off=off, seed=i, dt=1e-3))
        x = (np.arange(nn) - (nn - 1) / 2) * 0.5
        sig = rng.uniform(8, 25)
        envs.append(np.exp(-0.5 * (x[None, :] ** 2 + x[:, None] ** 2) / sig ** 2).ravel())
        off += nn * nn
    choice = tune_rk4(specs, "single", "double")
    tab = MemberTable(specs, choice.table_h)
    pump = DeviceArray.from_numpy(np.concatenate(envs).astype(np.float32))
    coef = DeviceArray.from_numpy(np.tile(rng.uniform(0.5, 3.0, 256) * pth, (4, 1)).astype(np.float32))
    e = Rk4Engine(tab, pump, "single", "double", whole_step=choice.whole_step, strip_h=choice.strip_h)
    e.set_dt(1e-3)
    e.psi[0].upload((1e-3 * (rng.standard_normal(off) + 1j * rng.standard_normal(off))).astype(np.complex64))
    e.run_steps(1000, coef, True, False, 0)
    e.stream.synchronize()
    start, stop = Event(), Event()
    start.record(e.stream)
    done = 0
    while done < steps:
        e.run_steps(1000, coef, True, False, done)
        done += 1000
    stop.record(e.stream)
    e.stream.synchronize()
    ms = start.elapsed_ms(stop)
    pt_steps = float(off) * steps
    return {"members": 256, "points": int(off), "steps": steps, "form": choice.form, "tile_h": choice.tile_h,
            "wall_ms": round(ms, 1), "point_steps_per_s": round(pt_steps / (ms * 1e-3), 1),
            "point_stage_updates_per_s": round(4 * pt_steps / (ms * 1e-3), 1),
            "effective_GBps_224B_model": round(224.0 * pt_steps / (ms * 1e-3) / 1e9, 1)}


def spinor_leg(n=1024, steps=20):
    """Spinor RK4 on one n^2 member, complex64, TE-TM 0.05, Zeeman 0.02, alpha_2 -1e-4, single
    reservoir per component, periodic grid: the whole-step kernel (one launch per step) against
    the four stage launches, each as graph-launched steps after one warm-up replay of the same
    graph (us per step)."""
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.device import DeviceArray, Event
    from polarism_amd.engine import Rk4Engine
    from polarism_amd.layout import MemberSpec, MemberTable
    p = PhysicsConstants()
    p.te_tm, p.zeeman, p.alpha_2 = 0.05, 0.02, -1e-4
    rng = np.random.default_rng(5)
    spec = MemberSpec(nx=n, ny=n, dx=0.3, dy=0.3, physics=p, reservoir="single", components=2, res_components=2,
                      dt=0.05, circular_polarization=0.3)
    pump = DeviceArray.from_numpy(rng.uniform(0.0, 1.0, n * n).astype(np.float32))
    coef = DeviceArray.from_numpy(np.full(4, 0.05, dtype=np.float32))
    psi = (0.3 * (rng.standard_normal(2 * n * n) + 1j * rng.standard_normal(2 * n * n))).astype(np.complex64)
    out = {"n": n, "components": 2, "res_components": 2}
    for name, whole in (("whole_step", True), ("stage_launches", False)):
        e = Rk4Engine(MemberTable([spec], 16), pump, "single", "single", whole_step=whole)
        e.set_dt(0.05)
        e.psi[0].upload(psi)
        e.run_steps(steps, coef, True, False, 0)
        start, stop = Event(), Event()
        start.record(e.stream)
        e.run_steps(steps, coef, True, False, 0)
        stop.record(e.stream)
        e.stream.synchronize()
        out[name + "_us_per_step"] = round(start.elapsed_ms(stop) * 1e3 / steps, 2)
    out["speedup"] = round(out["stage_launches_us_per_step"] / out["whole_step_us_per_step"], 3)
    return out


def quickstart_leg(steps=2000):
    """The reference quickstart on the GPU: rk4-fdm-hip, 128^2 periodic, lx = ly = 100 um,
    continuous Gaussian pump P0 0.02 / sigma 12 um, single reservoir, complex128, dt 1e-3 ps,
    2000 steps (2 ps) after one untimed pass that captures the graphs."""
    from polarism_amd import Config, SimulationController
    from polarism_amd.config import GridParameters
    from polarism_amd.device import Event
    cfg = Config()
    cfg.grid = GridParameters(nx=128, ny=128, lx=100.0, ly=100.0)
    cfg.laser.laser_type = "continuous-gaussian"
    cfg.laser.P0 = 0.02
    cfg.laser.sigma_space = 12.0
    cfg.reservoir.reservoir_type = "single"
    cfg.solver.method = "rk4-fdm-hip"
    cfg.solver.dt = 1e-3
    cfg.solver.total_time = steps * 1e-3
    cfg.solver.precision = "double"
    ctl = SimulationController(cfg)
    sol = ctl.solver
    sol.advance(ctl.model, ctl.state, steps)
    stream = sol.stream
    stream.synchronize()
    start, stop = Event(), Event()
    start.record(stream)
    sol.advance(ctl.model, ctl.state, steps)
    stop.record(stream)
    stream.synchronize()
    ms = start.elapsed_ms(stop)
    return {"steps": steps, "steps_per_s": round(steps / (ms * 1e-3), 1), "step_us": round(ms * 1e3 / steps, 2)}


def fft_leg(steps=1000):
    """Split-step FFT at 512^2 (complex64, Wigner noise), both operator orders."""
    from polarism_amd import Config, SimulationController
    from polarism_amd.config import GridParameters
    from polarism_amd.device import Event
    out = {}
    for method, transforms in (("split-step-fft-hip", 4), ("split-step-fft-kouter-hip", 2)):
        cfg = Config()
        cfg.grid = GridParameters(nx=512, ny=512, lx=200.0, ly=200.0)
        cfg.laser.laser_type = "continuous-gaussian"
        cfg.laser.sigma_space = 20.0
        cfg.laser.P0 = 2 * 0.02075
        cfg.reservoir.reservoir_type = "single"
        cfg.solver.method = method
        cfg.solver.dt = 5e-3
        cfg.solver.total_time = steps * 5e-3
        cfg.solver.precision = "single"
        cfg.noise.seed = 1234
        cfg.noise.stochastic = "wigner"
        ctl = SimulationController(cfg)
        sol = ctl.solver
        sol.advance(ctl.model, ctl.state, steps)     # captures the chunk's graph outside the timed region
        stream = sol.stream
        stream.synchronize()
        start, stop = Event(), Event()
        start.record(stream)
        sol.advance(ctl.model, ctl.state, steps)
        stop.record(stream)
        stream.synchronize()
        ms = start.elapsed_ms(stop)
        pts = 512 * 512
        out[method] = {"steps_per_s": round(steps / (ms * 1e-3), 1), "transforms_per_step": transforms,
                       "fft_GBps": round(transforms * 16.0 * pts * steps / (ms * 1e-3) / 1e9, 1),
                       "step_us": round(ms * 1e3 / steps, 2)}
    return out


def _oracle_rate(steps):
    """Point-steps/s of the oracle rk4-fdm step (complex64, one 256^2 member) on this core."""
    from oracle import gpe_oracle as O
    g = O.Grid(SWEEP["n"], SWEEP["n"], SWEEP["L"], SWEEP["L"])
    p = O.Physics()
    env = O.gaussian_profile(g, 12.0).astype(np.float32)
    pump_at = lambda t: env * np.float32(0.03)
    psi = O.initial_psi(g.shape, 1e-3, 0, "single").reshape(1, 1, *g.shape)
    res = (np.zeros((1, 1, *g.shape), np.float32),)
    psi, res = O.rk4_fdm_step("single", p, g, psi, res, pump_at, 0.0, SWEEP["dt"])
    t0 = time.perf_counter()
    for _ in range(steps):
        psi, res = O.rk4_fdm_step("single", p, g, psi, res, pump_at, 0.0, SWEEP["dt"])
    return steps * g.nx * g.ny / (time.perf_counter() - t0)


def _median_of(fn, repeats=3):
    vals = [fn() for _ in range(repeats)]
    return float(np.median(vals)), vals


def _oracle_configs():
    """The oracle (the reference's NumPy algorithms) on one core, each a median of three timings:
    the quickstart rk4-fdm (128^2 complex128, single reservoir), the split-step run with Wigner
    noise (512^2 complex64, CPU operator order) and the spinor Lawson DP5 trial (1024^2 complex64,
    double reservoir per component, ring pump)."""
    from oracle import gpe_oracle as O
    out = {}

    g = O.Grid(128, 128, 100.0, 100.0)
    p = O.Physics()
    env = O.gaussian_profile(g, 12.0)
    psi0 = O.initial_psi(g.shape, 1e-3, 0, "double").reshape(1, 1, *g.shape)

    def c1(steps=100):
        psi, res = psi0, (np.zeros((1, 1, *g.shape)),)
        t0 = time.perf_counter()
        for k in range(steps):
            psi, res = O.rk4_fdm_step("single", p, g, psi, res, lambda t: env * 0.02, k * 1e-3, 1e-3)
        return steps / (time.perf_counter() - t0)
    med, vals = _median_of(c1)
    out["quickstart_rk4_128_fp64"] = {"steps_per_s": round(med, 2), "repeats": [round(v, 2) for v in vals]}

    g2 = O.Grid(512, 512, 200.0, 200.0)
    env2 = O.gaussian_profile(g2, 20.0).astype(np.float32)
    psi2 = O.initial_psi(g2.shape, 1e-3, 1234, "single").reshape(1, 1, *g2.shape)
    nz = dict(seeds=[1234], gamma=[p.gamma_C], scattering=[p.R], strength=[1.0], volume=g2.dx * g2.dy,
              precision="single")

    def c2(steps=5):
        psi, res = psi2, (np.zeros((1, 1, *g2.shape), np.float32),)
        t0 = time.perf_counter()
        for k in range(steps):
            psi, res = O.split_step_cpu("single", p, g2, psi, res, lambda t: env2 * np.float32(0.0415),
                                        k * 5e-3, 5e-3, k, nz)
        return steps / (time.perf_counter() - t0)
    med, vals = _median_of(c2)
    out["split_step_512_fp32_noise"] = {"steps_per_s": round(med, 3), "repeats": [round(v, 3) for v in vals]}

    g3 = O.Grid(1024, 1024, 300.0, 300.0)
    p3 = O.Physics(te_tm=0.05, zeeman=0.02, alpha_2=-0.1 * O.Physics().g_C)
    env3 = O.ring_profile(g3, 15.0, 3.0).astype(np.float32)
    weights = np.asarray([0.75, 0.25], dtype=np.float32).reshape(1, 2, 1, 1)
    prob = O.LawsonProblem("double", p3, g3, precision="single", tableau="dp5", adaptive=True, rtol=1e-5,
                           atol=1e-8, pump_weights=weights, reservoir_components=2)
    psi3 = np.stack((O.initial_psi(g3.shape, 1e-3, 3, "single"), O.initial_psi(g3.shape, 1e-3, 4, "single")))[None]
    psi3k = np.fft.fft2(psi3, axes=(-2, -1)).astype(np.complex64)
    res3 = tuple(np.zeros((1, 2) + g3.shape, np.float32) for _ in range(2))

    def c3(trials=2):
        t0 = time.perf_counter()
        for _ in range(trials):
            prob.step(psi3k, res3, lambda t: env3 * np.float32(0.1), 0.0, 0.05)
        return trials / (time.perf_counter() - t0)
    med, vals = _median_of(c3)
    out["lawson_dp5_spinor_1024_fp32"] = {"trials_per_s": round(med, 4), "repeats": [round(v, 4) for v in vals]}
    return out


def cpu_worker(steps, procs):
    """Runs in a child process that never touches the GPU: the sweep's rk4-fdm member step on one
    core and on a pool of ``procs`` (one member per process), each a median of three timings, then
    the other configurations on one core."""
    import multiprocessing as mp
    single, singles = _median_of(lambda: _oracle_rate(steps))
    progress("cpu: one-core rk4 step timed")
    pooled_runs = []
    with mp.get_context("fork").Pool(procs) as pool:
        for k in range(3):
            t0 = time.perf_counter()
            pool.map(_oracle_rate, [steps] * procs)
            pooled_runs.append(procs * steps * SWEEP["n"] ** 2 / (time.perf_counter() - t0))
            progress(f"cpu: pooled run {k + 1}/3 timed")
    configs = _oracle_configs()
    progress("cpu: configurations timed")
    print(json.dumps({"single": single, "singles": singles, "pooled": float(np.median(pooled_runs)),
                      "pooled_runs": pooled_runs, "procs": procs, "configs": configs}))


def cpu_baseline(point_steps, converged, steps=200):
    """The oracle rk4-fdm (the reference's NumPy algorithm) on the host cores, in a child process:
    one member per core, scaled to the plain sweep's point-steps."""
    import subprocess
    procs = max(1, min(16, len(os.sched_getaffinity(0))))
    # the worker's progress lines go straight to this process's stderr
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", str(steps), str(procs)],
                         stdout=subprocess.PIPE, text=True, timeout=600, cwd=ROOT)
    if out.returncode != 0:
        raise RuntimeError(f"cpu worker failed with status {out.returncode}")
    r = json.loads(out.stdout.strip().splitlines()[-1])
    pooled, single = r["pooled"], r["single"]
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": round(converged / (point_steps / pooled), 6), "unit": "members/s", "cores": procs,
            "kind": "port",
            "sample": f"oracle rk4-fdm (NumPy restatement of rk4_fdm_solver.py), complex64 256x256, {steps} steps "
                      f"per process on {procs} processes ({pooled / 1e6:.1f} Mpt-step/s pooled, median of 3; "
                      f"{single / 1e6:.2f} on one core, median of 3), scaled to the plain sweep's "
                      f"{point_steps:.3e} point-steps",
            "pooled_mpt_step_per_s_runs": [round(v / 1e6, 2) for v in r["pooled_runs"]],
            "single_core_mpt_step_per_s_runs": [round(v / 1e6, 3) for v in r["singles"]],
            "single_core_members_per_s": round(converged / (point_steps / single), 6),
            "configs_one_core": r["configs"],
            "cpu_model": cpu_model, "affinity_cores": len(os.sched_getaffinity(0))}


def fft_hbm_leg(n=2048, planes=16, reps=10):
    """Batched forward 2-D FFT of ``planes`` complex64 n x n planes (n = 2048, 16 planes: 537 MB,
    beyond the Infinity Cache): FFT GB/s with 2 x 8 B per point per 2-D transform (one read and
    one write, the single-pass ideal) and the HBM bytes of the two-pass (rows, columns) kernels."""
    import ctypes as C
    from polarism_amd import _native as N
    from polarism_amd.device import DeviceArray, Event, Stream
    from polarism_amd.solvers.split_step_hip import FFTPlan
    s = Stream()
    plan = FFTPlan(n, n, 0.5, 0.5, "single")
    rng = np.random.default_rng(1)
    x = DeviceArray.from_numpy((rng.standard_normal(n * n * planes) + 0j).astype(np.complex64))
    y = DeviceArray.empty((n * n * planes,), np.complex64)
    run = lambda: N.call("pgpe_fft2", plan.handle, x.ptr, y.ptr, planes, 0, s.handle)
    run()
    s.synchronize()
    start, stop = Event(), Event()
    start.record(s)
    for _ in range(reps):
        run()
    stop.record(s)
    s.synchronize()
    ms = start.elapsed_ms(stop) / reps
    pts = float(n) * n * planes
    ideal = 16.0 * pts / (ms * 1e-3) / 1e9
    return {"n": n, "planes": planes, "transform_us": round(ms * 1e3, 1), "fft_GBps_single_pass_model": round(ideal, 1),
            "frac_of_hbm_peak": round(ideal / HBM_PEAK_GBS, 4), "two_pass_GBps": round(2 * ideal, 1)}


def lawson_leg(ticks=20):
    """Spinor Lawson IFRK, DP5 adaptive (1 ps ticks), 1024^2, lx = ly = 300 um, TE-TM 0.05,
    Zeeman 0.02, alpha_2 = -0.1 g_C, elliptic ring pump (r 15 um, width 3 um), double reservoir,
    complex64 fields with FP64 error norms: a window of ``ticks`` ps after one warm-up tick, with
    the trial controller on the device (captured trial graphs) and on the host, advancing one
    tick per call and all ``ticks`` in one call (which lands only on the last tick).
    rtol is 1e-5, the smallest the reference's validation accepts in single precision
    (validation.py:241-242); 1e-6 is accepted in double precision only."""
    from polarism_amd import Config, SimulationController
    from polarism_amd.config import GridParameters
    from polarism_amd.device import Event
    cfg = Config()
    cfg.grid = GridParameters(nx=1024, ny=1024, lx=300.0, ly=300.0)
    cfg.physics.spin = "spinor"
    cfg.physics.te_tm = 0.05
    cfg.physics.zeeman = 0.02
    cfg.physics.alpha_2 = -0.1 * cfg.physics.g_C
    cfg.laser.laser_type = "profile"
    cfg.laser.profile = {"type": "ring", "radius": 15.0, "width": 3.0}
    cfg.laser.P0 = 0.1
    cfg.laser.circular_polarization = 0.5
    cfg.reservoir.reservoir_type = "double"
    cfg.solver.method = "ifrk-fft-hip"
    cfg.solver.tableau = "dp5"
    cfg.solver.adaptive = True
    cfg.solver.rtol = 1e-5
    cfg.solver.dt = 1.0
    cfg.solver.total_time = float(ticks + 1)
    cfg.solver.precision = "single"
    cfg.noise.seed = 3
    out = {}
    pts = 1024 * 1024 * 2
    # per trial: 7 stages, each a forward and an inverse 2-D transform (2 passes each), plus
    # the final combination's inverse (+ the error field's inverse) -> ~ (2 * 7 + 2) transforms
    transforms = 2 * 7 + 2
    for device in (True, False):
        for chunk in (1, ticks):
            ctl = SimulationController(cfg)
            sol = ctl.solver
            sol.device_trials = device
            sol.advance(ctl.model, ctl.state, 1)
            sc = sol.step_controller
            a0, r0 = sc.accepted, sc.rejected
            sol.stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(ticks // chunk):
                sol.advance(ctl.model, ctl.state, chunk)
            sol.stream.synchronize()
            wall = time.perf_counter() - t0
            trials = (sc.accepted - a0) + (sc.rejected - r0)
            name = ("device" if device else "host") + ("_controller" if chunk == 1 else f"_controller_advance_{chunk}")
            out[name] = {
                "ticks": ticks, "ticks_per_advance": chunk, "rtol": 1e-5, "trials": trials,
                "accepted": sc.accepted - a0, "rejected": sc.rejected - r0,
                "ms_per_trial": round(wall * 1e3 / max(trials, 1), 3), "ps_per_s": round(ticks / wall, 3),
                "dt_last_ps": sc.dt_last,
                "fft_GBps_16_transforms_per_trial": round(transforms * 16.0 * pts * trials / wall / 1e9, 1)}
    # the complex64 run against the same configuration in complex128 after a few ticks (each with
    # its own adaptive trial sequence at rtol 1e-5, so the difference holds both the precision
    # and the step choices)
    span = 6
    fields, counts = {}, {}
    for prec in ("single", "double"):
        cfg.solver.precision = prec
        cfg.solver.total_time = float(span)
        ctl = SimulationController(cfg)
        ctl.solver.advance(ctl.model, ctl.state, span)
        ctl.solver.stream.synchronize()
        fields[prec] = ctl.state.field.get().astype(np.complex128)
        sc = ctl.solver.step_controller
        counts[prec] = {"accepted": sc.accepted, "rejected": sc.rejected}
    a, b = fields["single"], fields["double"]
    rho_a, rho_b = np.abs(a) ** 2, np.abs(b) ** 2
    out["fp32_vs_fp64"] = {
        "ps": span, "trials": counts,
        "psi_rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
        "density_rel_l2": float(np.linalg.norm(rho_a - rho_b) / np.linalg.norm(rho_b)),
        "norm_rel_diff": float(abs(rho_a.sum() - rho_b.sum()) / rho_b.sum())}
    cfg.solver.precision = "single"
    return out


def fno_leg(n=8192, reps=10):
    """The surrogate's MFMA kernel and network: a square BF16 GEMM (n^3, FP32 accumulate) in
    TFLOP/s against the 2.5 PFLOP/s dense BF16 peak, and the FNO's forward pass and one training
    step (forward, hand-written backward, AdamW) at batch 16 on 256^2 maps."""
    import torch
    from polarism_amd.surrogate.fno import FNOConfig, FNOModule
    from polarism_amd.surrogate.kernels import gemm
    dev = torch.device("cuda")
    a = torch.randn(n, n, device=dev).to(torch.bfloat16)
    b = torch.randn(n, n, device=dev).to(torch.bfloat16)
    c = torch.empty(n, n, device=dev)
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def gemm_ms(sb):
        run = lambda: gemm(a, b, c, n, n, n, sa=(n, 1, 0), sb=sb, sc=(n, 1, 0))
        run()
        torch.cuda.synchronize()
        start.record()
        for _ in range(reps):
            run()
        stop.record()
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / reps

    # B row-major (n contiguous), then B with k contiguous (the transposed operand)
    ms = gemm_ms((n, 1, 0))
    tflops = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
    ms_t = gemm_ms((1, n, 0))
    tflops_t = 2.0 * n ** 3 / (ms_t * 1e-3) / 1e12
    cfg = FNOConfig()
    mod = FNOModule(cfg, seed=0)
    x = torch.randn(cfg.fin, 16, cfg.H, cfg.W, device=dev)
    y = torch.randn(cfg.fout, 16, cfg.H, cfg.W, device=dev)
    opt = torch.optim.AdamW(mod.parameters(), lr=1e-4)

    def train_step():
        opt.zero_grad(set_to_none=True)
        loss = ((mod(x) - y) ** 2).mean()
        loss.backward()
        opt.step()

    for fn in (train_step,):
        fn()
    torch.cuda.synchronize()
    t = {}
    for name, fn in (("forward", lambda: mod(x)), ("train_step", train_step)):
        start.record()
        for _ in range(reps):
            with torch.no_grad() if name == "forward" else torch.enable_grad():
                fn()
        stop.record()
        torch.cuda.synchronize()
        t[name] = start.elapsed_time(stop) / reps
    return {"gemm_bf16": {"n": n, "avg_us": round(ms * 1e3, 1), "TFLOPs": round(tflops, 1),
                          "frac_of_2500_TFLOPs_dense_bf16": round(tflops / 2500.0, 4)},
            "gemm_bf16_k_contiguous_b": {"n": n, "avg_us": round(ms_t * 1e3, 1), "TFLOPs": round(tflops_t, 1),
                                         "frac_of_2500_TFLOPs_dense_bf16": round(tflops_t / 2500.0, 4)},
            "fno_batch16_256": {"params": sum(p.numel() for p in mod.parameters()),
                                "forward_ms": round(t["forward"], 2), "train_step_ms": round(t["train_step"], 2)}}


GATE = 1e-3


def density_error(a, b):
    """Relative L2 of density |a|^2 against the judge's |b|^2 (two zero fields agree)."""
    da, db = np.abs(a) ** 2, np.abs(b) ** 2
    den = np.linalg.norm(db)
    if den < 1e-30 and np.linalg.norm(da) < 1e-30:
        return 0.0
    return float(np.linalg.norm(da - db) / max(den, 1e-30))


def gated_rates(cold, warm, infer_s, zero):
    """Members/s of both forms on the gated set: plain converged, warm converged, densities within GATE."""
    cold_ok, warm_ok = np.isfinite(cold.stop_times), np.isfinite(warm.stop_times)
    both = cold_ok & warm_ok
    err = np.full(cold_ok.size, np.nan)
    for i in np.flatnonzero(both):
        err[i] = density_error(warm.final_psi[i], cold.final_psi[i])
    passed = both & (err <= GATE)

    def rate(res, extra):
        if not passed.any():
            return 0.0, None
        wall = float(np.max(res.wall_at_stop[passed])) + extra
        return int(passed.sum()) / wall, wall

    warm_rate, warm_wall = rate(warm, infer_s)
    plain_rate, plain_wall = rate(cold, 0.0)
    plain_own = cold_ok.sum() / cold.wall_to_last_s if cold_ok.any() else 0.0
    e = err[both]
    return {
        "gated_members": int(passed.sum()), "members": int(cold_ok.size),
        "warm_members_per_s": round(warm_rate, 4), "plain_members_per_s": round(plain_rate, 4),
        "speedup_warm_over_plain": round(warm_rate / plain_rate, 3) if plain_rate > 0 else None,
        "warm_wall_to_last_gated_s": None if warm_wall is None else round(warm_wall, 3),
        "plain_wall_to_last_gated_s": None if plain_wall is None else round(plain_wall, 3),
        "inference_s": round(infer_s, 3),
        "converged_plain": int(cold_ok.sum()), "converged_warm": int(warm_ok.sum()),
        "compared": int(both.sum()), "pass_1e-3": int(passed.sum()),
        "failed_gate": int((both & ~passed).sum()), "warm_only_converged": int((warm_ok & ~cold_ok).sum()),
        "median_density_rel_l2": float(np.median(e)) if e.size else None,
        "max_density_rel_l2": float(np.max(e)) if e.size else None,
        "max_density_rel_l2_gated": float(np.max(err[passed])) if passed.any() else None,
        "zero_starts": int(zero.sum()), "zero_starts_gated": int((zero & passed).sum()),
        "zero_starts_cold_condensed": int((zero & both & ~passed).sum()),
        "zero_starts_cold_unconverged": int((zero & ~cold_ok).sum()),
        "median_stop_ps_gated": {"plain": float(np.median(cold.stop_times[passed])) if passed.any() else None,
                                 "warm": float(np.median(warm.stop_times[passed])) if passed.any() else None},
        "plain_own_members_per_s": round(float(plain_own), 4),
    }, passed, err


def sensitivity_sample(members, cold, passed, count=8, window=200.0):
    """Is the plain FP32 steady state (dt 0.05 ps) also a steady state in complex128 and at dt / 2?
    ``count`` gated members that hold a condensate restart from their final FP32 state and run the
    same criterion for up to ``window`` ps; reported: how many meet it and how far their density
    moved.  (A cold start at dt / 2 converges much more slowly: RK4's damping of the stiffest
    Laplacian modes, which speeds the approach at dt 0.05 ps, is far weaker there.)"""
    condensed = np.asarray([(np.abs(p) ** 2).max() > 0.05 for p in cold.final_psi])
    idx = np.flatnonzero(passed & condensed)[:count]
    if idx.size == 0:
        return None
    from polarism_amd.sweep import PackedSweep, SweepMember
    sub = [SweepMember(members[i].nx, members[i].ny, members[i].lx, members[i].ly, members[i].physics,
                       members[i].envelope, members[i].P0, seed=members[i].seed, psi0=cold.final_psi[i],
                       res0=cold.final_res[i]) for i in idx]
    out = {"members": [int(i) for i in idx], "window_ps": window}
    for name, prec, dt in (("fp64", "double", SWEEP["dt"]), ("half_dt", "single", SWEEP["dt"] / 2)):
        sw = PackedSweep(sub, reservoir="single", precision=prec, dt=dt, check_every=SWEEP["check_every"],
                         rtol=SWEEP["rtol"], max_time=window)
        res = sw.run()
        errs = [density_error(res.final_psi[k], cold.final_psi[i]) for k, i in enumerate(idx)]
        out[name] = {"converged": int(np.isfinite(res.stop_times).sum()),
                     "median_density_rel_l2": float(np.median(errs)), "max_density_rel_l2": float(np.max(errs))}
    return out


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--cpu-worker":
        cpu_worker(int(sys.argv[2]), int(sys.argv[3]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--members", type=int, default=SWEEP["members"])
    ap.add_argument("--legs", default="sweep,copy,roofline,uneven,quickstart,spinor,fft,ffthbm,lawson,fno,cpu")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--dump", default=None, help="write per-member stop times and density errors to this .npz")
    ap.add_argument("--max-time", type=float, default=SWEEP["max_time"],
                    help="sweep horizon in ps (the workload's is 4000; shorter only to bound profiler traces)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the sweep's steps one by one instead of as captured graph chunks (same "
                         "kernels; for kernel-trace profilers that cannot follow graph launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (the sweep would shard "
              "members across GPUs; no multi-GPU run is part of it)", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))
    from polarism_amd.device import init
    from polarism_amd.surrogate import warmstart as ws
    init()
    result = {"metric": "steady-state sweep members/s (warm vs plain) at <=1e-3 rel L2; % HBM roofline",
              "unit": "members/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
              "data": "synthetic (seeded random pumps, noise-seeded initial fields; surrogate trained on "
                      "solver-generated steady states)"}
    SWEEP["max_time"] = args.max_time
    members = ws.make_members(args.members, SWEEP["seed"])
    have_weights = os.path.exists(args.weights)
    if "sweep" in legs:
        for k in range(args.warmup):
            run_sweep(ws.make_members(8, SWEEP["seed"] + 1), graphs=not args.no_graphs)
            progress(f"warmup {k + 1}/{args.warmup} done")
        plain_runs, warm_runs, infer_s = [], [], []
        t_all = time.perf_counter()
        for k in range(args.steps):
            t0 = time.perf_counter()
            plain_runs.append(run_sweep(members, graphs=not args.no_graphs))
            line = f"step {k + 1}/{args.steps}: plain {time.perf_counter() - t0:.2f} s"
            if have_weights:
                params, cfg, norm = ws.load_weights(args.weights)
                t0 = time.perf_counter()
                wm = ws.warm_start(members, None, norm, states=ws.predict_states(params, members, cfg, norm))
                infer_s.append(time.perf_counter() - t0)
                warm_runs.append(run_sweep(wm, graphs=not args.no_graphs))
                line += f", warm {time.perf_counter() - t0:.2f} s"
            progress(line)
        wall = time.perf_counter() - t_all
        plain = summarise(plain_runs[-1])
        sweep = {"plain": plain, "warm": None}
        value = plain["members_per_s"]
        if have_weights:
            zero = np.asarray([m.label.get("warm") == "zero" for m in wm])
            gate, passed, err = gated_rates(plain_runs[-1], warm_runs[-1], infer_s[-1], zero)
            sweep["warm"] = summarise(warm_runs[-1], extra_s=infer_s[-1])
            sweep["gated"] = gate
            sweep["speedup_warm_over_plain"] = gate["speedup_warm_over_plain"]
            value = gate["warm_members_per_s"]
            sweep["sensitivity"] = sensitivity_sample(members, plain_runs[-1], passed)
            if args.dump:
                c, w = plain_runs[-1], warm_runs[-1]
                np.savez(args.dump, cold_stop=c.stop_times, warm_stop=w.stop_times, cold_wall=c.wall_at_stop,
                         warm_wall=w.wall_at_stop, err=err, zero=zero,
                         ratio=np.asarray([m.label.get("ratio", np.nan) for m in members]),
                         kind=np.asarray([m.label.get("kind", "") for m in members]))
        result.update(value=value, ms_per_step=round(wall / args.steps * 1e3, 1))
        result["sweep"] = sweep
    if "copy" in legs:
        result["device_copy"] = copy_leg()
        progress("copy leg done")
    if "roofline" in legs:
        result["roofline"] = roofline_leg()
        progress("roofline leg done")
    if "uneven" in legs:
        result["uneven_grid_sweep"] = uneven_leg()
        progress("uneven leg done")
    if "quickstart" in legs:
        result["quickstart_rk4_fp64"] = quickstart_leg()
        progress("quickstart leg done")
    if "spinor" in legs:
        result["spinor_rk4_1024"] = spinor_leg()
        progress("spinor leg done")
    if "fft" in legs:
        result["split_step_fft"] = fft_leg()
        progress("fft leg done")
    if "ffthbm" in legs:
        result["fft_batched_2048"] = fft_hbm_leg()
        progress("ffthbm leg done")
    if "lawson" in legs:
        result["lawson_spinor_dp5"] = lawson_leg()
        progress("lawson leg done")
    if "fno" in legs:
        result["surrogate_fno"] = fno_leg()
        progress("fno leg done")
    result["config"] = {"workload": "steady-state pump sweep, 256 members of 256x256, rk4-fdm-hip complex64, "
                                    "single reservoir, dt 0.05 ps; warm start (FNO) vs plain cold start",
                        "members": len(members), "grid": "256x256", "lx_um": SWEEP["L"], "dt_ps": SWEEP["dt"],
                        "steady_rtol": SWEEP["rtol"], "check_every_ps": SWEEP["check_every"],
                        "max_time_ps": SWEEP["max_time"], "surrogate": os.path.basename(args.weights)
                        if have_weights else None, "parallelism": "1 GPU, members packed in one launch"}
    if "cpu" in legs and "sweep" in result:
        s = result["sweep"]["plain"]
        result["cpu_baseline"] = cpu_baseline(s["point_steps"], max(s["converged"], 1))
        cpu, ratios = result["cpu_baseline"]["configs_one_core"], {}
        if "quickstart_rk4_fp64" in result:
            ratios["quickstart_rk4_128_fp64"] = round(result["quickstart_rk4_fp64"]["steps_per_s"]
                                                      / cpu["quickstart_rk4_128_fp64"]["steps_per_s"], 1)
        if "split_step_fft" in result:
            ratios["split_step_512_fp32_noise"] = round(result["split_step_fft"]["split-step-fft-hip"]["steps_per_s"]
                                                        / cpu["split_step_512_fp32_noise"]["steps_per_s"], 1)
        if "lawson_spinor_dp5" in result:
            ratios["lawson_dp5_spinor_1024_fp32"] = round(
                1e3 / result["lawson_spinor_dp5"]["device_controller"]["ms_per_trial"]
                / cpu["lawson_dp5_spinor_1024_fp32"]["trials_per_s"], 1)
        result["cpu_baseline"]["gpu_over_one_core"] = ratios
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
