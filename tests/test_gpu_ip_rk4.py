"""ip-rk4-hip against the oracle's interaction-picture RK4 (solver/ip_rk4_solver.py:122-213)."""
import numpy as np
import pytest

from oracle import gpe_oracle as O
from tests.helpers import oracle_physics, quickstart_config

pytestmark = pytest.mark.gpu


def _oracle(cfg, n_steps, spinor):
    g = O.Grid(cfg.grid.nx, cfg.grid.ny, cfg.grid.lx, cfg.grid.ly)
    prec = cfg.solver.precision
    real = np.float32 if prec == "single" else np.float64
    kind = cfg.reservoir.reservoir_type
    p = oracle_physics(cfg)
    psi = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed, prec)
    if spinor:
        other = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed + 1, prec)
        psi = np.stack((psi, other))[None]
    else:
        psi = psi.reshape(1, 1, *g.shape)
    C_res = 2 if spinor else 1
    nf = {"none": 0, "single": 1, "double": 2, "quadratic-double": 2}[kind]
    res = tuple(np.zeros((1, C_res, *g.shape), dtype=real) for _ in range(nf))
    env = O.gaussian_profile(g, cfg.laser.sigma_space).astype(real)
    pump_at = lambda t: env * cfg.laser.P0
    weights = None
    if spinor:
        cp = cfg.laser.circular_polarization
        weights = np.asarray([(1 + cp) / 2, (1 - cp) / 2], dtype=real).reshape(1, 2, 1, 1)
    prob = O.LawsonProblem(kind, p, g, precision=prec, tableau="rk4", pump_weights=weights,
                           reservoir_components=C_res)
    dt = cfg.solver.dt
    for s in range(n_steps):
        psi, res = O.ip_rk4_step(prob, psi, res, pump_at, s * dt, dt)
    return psi, res


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("kind", ["single", "double", "quadratic-double"])
def test_scalar_ip_rk4_matches_oracle(precision, kind):
    """64 x 96 grid, Gaussian pump above threshold, kinetic relaxation on: 10 steps, rel-L2 on psi
    and every reservoir field <= 1e-10 (FP64) / 1e-5 (FP32)."""
    cfg = quickstart_config(solver__method="ip-rk4-hip", grid__nx=64, grid__ny=96, grid__lx=40.0, grid__ly=60.0,
                            solver__dt=0.02, solver__total_time=0.2, laser__P0=0.3, laser__sigma_space=8.0,
                            reservoir__reservoir_type=kind, solver__precision=precision,
                            physics__kinetic_relaxation_eta=0.03)
    from polarism_amd import SimulationController
    ctl = SimulationController(cfg)
    ctl.run()
    psi_ref, res_ref = _oracle(cfg, 10, spinor=False)
    tol = 1e-10 if precision == "double" else 1e-5
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= tol
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= tol


@pytest.mark.parametrize("reservoir", ["single", "double"])
def test_spinor_ip_rk4_matches_oracle(reservoir):
    cfg = quickstart_config(solver__method="ip-rk4-hip", grid__nx=64, grid__ny=64, grid__lx=40.0, grid__ly=40.0,
                            solver__dt=0.02, solver__total_time=0.2, laser__P0=0.3, laser__sigma_space=8.0,
                            reservoir__reservoir_type=reservoir, physics__spin="spinor", physics__te_tm=0.05,
                            physics__zeeman=0.02, physics__alpha_2=-1e-4, laser__circular_polarization=0.3)
    from polarism_amd import SimulationController
    ctl = SimulationController(cfg)
    ctl.run()
    psi_ref, res_ref = _oracle(cfg, 10, spinor=True)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= 1e-10
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= 1e-10
