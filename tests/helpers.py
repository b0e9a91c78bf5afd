"""Shared test constructors: configs of the reference's workloads and oracle runners."""
from __future__ import annotations

import numpy as np

from oracle import gpe_oracle as O
from polarism_amd.config import Config, GridParameters


def quickstart_config(**overrides) -> Config:
    """The reference quickstart: 128^2 periodic, lx = ly = 100 um, continuous Gaussian
    P0 = 0.02, sigma 12 um, dt 1e-3 ps, 2 ps, single reservoir, complex128, seed 0."""
    cfg = Config()
    cfg.grid = GridParameters(nx=128, ny=128, lx=100.0, ly=100.0)
    cfg.laser.laser_type = "continuous-gaussian"
    cfg.laser.P0 = 0.02
    cfg.laser.sigma_space = 12.0
    cfg.reservoir.reservoir_type = "single"
    cfg.solver.method = "rk4-fdm-hip"
    cfg.solver.dt = 1e-3
    cfg.solver.total_time = 2.0
    cfg.solver.precision = "double"
    cfg.noise.seed = 0
    cfg.noise.initial_amplitude = 1e-3
    for path, value in overrides.items():
        section, name = path.split("__", 1)
        setattr(getattr(cfg, section), name, value)
    return cfg


def oracle_physics(cfg) -> O.Physics:
    p = cfg.physics
    return O.Physics(hbar=p.hbar, m_eff=p.m_eff, gamma_R=p.gamma_R, gamma_C=p.gamma_C, g_C=p.g_C, g_R=p.g_R,
                     g_I=p.g_I, gamma_I=p.gamma_I, gamma_A=p.gamma_A, R=p.R, R_IA=p.R_IA, R_AI=p.R_AI,
                     kappa=p.kappa, kinetic_relaxation_eta=p.kinetic_relaxation_eta, alpha_2=p.alpha_2,
                     zeeman=p.zeeman, te_tm=p.te_tm)


def oracle_rk4_run(cfg, n_steps, psi0=None, res0=None, noise=False, potential=0.0):
    """Run the oracle rk4-fdm on a single-member config; returns (psi (1,1,ny,nx), res tuple).

    The grid type and ``solver.laplacian`` come from the config."""
    g = O.Grid(cfg.grid.nx, cfg.grid.ny, cfg.grid.lx, cfg.grid.ly, boundary=cfg.grid.grid_type)
    kw = dict(potential=potential, laplacian_kind=cfg.solver.laplacian)
    prec = cfg.solver.precision
    real = np.float32 if prec == "single" else np.float64
    kind = cfg.reservoir.reservoir_type
    p = oracle_physics(cfg)
    if psi0 is None:
        psi0 = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed, prec).reshape(1, 1, *g.shape)
    nf = {"none": 0, "single": 1, "double": 2, "quadratic-double": 2}[kind]
    if res0 is None:
        res0 = tuple(np.zeros((1, 1, *g.shape), dtype=real) for _ in range(nf))
    env = O.gaussian_profile(g, cfg.laser.sigma_space, cfg.laser.x0, cfg.laser.y0).astype(real)
    if cfg.laser.laser_type == "pulse-gaussian":
        L = cfg.laser
        coef = lambda t: O.pulse_gaussian_coeff(t, L.P0, L.Pmax, L.sigma_time, L.cutoff_sigma,
                                                L.pulse_separation, L.n_pulses, L.delay)
    else:
        coef = lambda t: cfg.laser.P0 if t >= cfg.laser.delay else 0.0
    pump_at = lambda t: env * coef(t)
    nz = None
    if noise:
        nz = dict(seeds=[cfg.noise.seed], gamma=[p.gamma_C], scattering=[p.R],
                  strength=[cfg.noise.stochastic_strength], volume=g.dx * g.dy, precision=prec)
    psi, res = psi0, res0
    dt = cfg.solver.dt
    for s in range(n_steps):
        psi, res = O.rk4_fdm_step(kind, p, g, psi, res, pump_at, s * dt, dt, **kw)
        if nz is not None:
            O.wigner_add(psi, res[0] if res else 0, dt=dt, step=s, **nz)
    return psi, res
