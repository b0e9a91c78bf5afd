"""ifrk-fft-hip (Lawson IFRK) against the reference's golden trajectories and the oracle."""
from pathlib import Path

import numpy as np
import pytest

from oracle import gpe_oracle as O
from tests.helpers import oracle_physics, quickstart_config
from tests.test_oracle_golden import GOLDEN, fixture_case

pytestmark = pytest.mark.gpu


def _ctl(cfg):
    from polarism_amd import SimulationController
    return SimulationController(cfg)


def small_cfg(**kw):
    base = dict(grid__nx=16, grid__ny=16, grid__lx=20.0, grid__ly=20.0, solver__method="ifrk4-fft-hip",
                solver__dt=0.007, solver__total_time=0.035, laser__laser_type="pulse-gaussian",
                laser__P0=0.4, laser__sigma_space=15.0)
    base.update(kw)
    cfg = quickstart_config(**base)
    cfg.laser.sigma_time = 0.03
    cfg.laser.cutoff_sigma = 3.0
    return cfg


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("kind", ["single", "quadratic-double"])
@pytest.mark.parametrize("eta", [0.0, 0.04])
def test_hip_lawson_rk4_matches_reference_golden_trajectories(precision, kind, eta):
    """The HIP path itself against the reference's reference.npz (tolerances of test_ifrk_reference.py:19-21)."""
    g, p, psi, res, pot, _ = fixture_case(precision, kind, eta)
    cfg = small_cfg(solver__precision=precision, reservoir__reservoir_type=kind,
                    physics__kinetic_relaxation_eta=eta)
    ctl = _ctl(cfg)
    ctl.state.psi = psi[0, 0]
    ctl.reservoir.set_state(res)
    ctl.model.potential = pot
    for _ in range(5):
        ctl.solver.step(ctl.model, ctl.state, cfg.solver.dt)
    key = f"{precision}_{kind}_eta{eta:g}"
    tol = 1e-5 if precision == "single" else 1e-12
    with np.load(GOLDEN) as ref:
        r = ref[key + "_psi"]
        np.testing.assert_allclose(ctl.state.psi, r, rtol=tol, atol=tol * max(1e-3, np.max(abs(r))))
        for i, field in enumerate(ctl.reservoir.get_state()):
            r = ref[f"{key}_res{i}"]
            np.testing.assert_allclose(field.get()[0, 0], r, rtol=tol, atol=tol * max(1e-3, np.max(abs(r))))


def _oracle_fixed(cfg, n_steps, tableau, spinor=False):
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
    prob = O.LawsonProblem(kind, p, g, precision=prec, tableau=tableau, pump_weights=weights,
                           reservoir_components=C_res)
    dt = cfg.solver.dt
    for s in range(n_steps):
        _, _, psi, res = prob.step(psi, res, pump_at, s * dt, dt)
    return psi, res


@pytest.mark.parametrize("tableau", ["rk4", "dp5", "dop853"])
@pytest.mark.parametrize("kind", ["single", "double"])
def test_fixed_step_tableaus_match_oracle(tableau, kind):
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__tableau=tableau, grid__nx=64, grid__ny=96,
                            grid__lx=40.0, grid__ly=60.0, solver__dt=0.02, solver__total_time=0.2,
                            laser__P0=0.3, laser__sigma_space=8.0, reservoir__reservoir_type=kind)
    ctl = _ctl(cfg)
    ctl.run()
    psi_ref, res_ref = _oracle_fixed(cfg, 10, tableau)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= 1e-10
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= 1e-10


@pytest.mark.parametrize("tableau", ["rk4", "dp5"])
@pytest.mark.parametrize("reservoir", ["single", "double"])
def test_spinor_fixed_step_matches_oracle(tableau, reservoir):
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__tableau=tableau, grid__nx=64, grid__ny=64,
                            grid__lx=40.0, grid__ly=40.0, solver__dt=0.02, solver__total_time=0.2,
                            laser__P0=0.3, laser__sigma_space=8.0, reservoir__reservoir_type=reservoir,
                            physics__spin="spinor", physics__te_tm=0.05, physics__zeeman=0.02,
                            physics__alpha_2=-1e-4, laser__circular_polarization=0.3)
    ctl = _ctl(cfg)
    ctl.run()
    psi_ref, res_ref = _oracle_fixed(cfg, 10, tableau, spinor=True)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= 1e-10
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= 1e-10


def _uniform_cfg(tableau, rtol, dt=0.5, total=2.0, precision="double"):
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__tableau=tableau, grid__nx=16, grid__ny=16,
                            grid__lx=20.0, grid__ly=20.0, solver__dt=dt, solver__total_time=total,
                            solver__adaptive=True, solver__rtol=rtol, solver__atol=rtol * 0.01,
                            reservoir__reservoir_type="none", noise__enabled=False, laser__P0=0.0,
                            physics__g_C=2.0, physics__gamma_C=0.7, solver__precision=precision)
    return cfg


def _exact_uniform(cfg):
    t = cfg.solver.total_time
    gamma = cfg.physics.gamma_C
    phase = -cfg.physics.g_C / (cfg.physics.hbar * gamma) * (1 - np.exp(-gamma * t))
    return np.exp(-gamma * t / 2 + 1j * phase)


@pytest.mark.parametrize("tableau", ["dp5", "dop853"])
def test_adaptive_uniform_decay_closed_form(tableau):
    """tests/unit/test_adaptive_dt.py:34-80: error <= 50 rtol; tighter rtol takes more steps."""
    accepted = []
    for rtol in (1e-6, 1e-9):
        cfg = _uniform_cfg(tableau, rtol)
        ctl = _ctl(cfg)
        ctl.state.psi = np.ones((16, 16), dtype=complex)
        ctl.run()
        err = np.max(np.abs(ctl.state.psi - _exact_uniform(cfg)))
        assert err <= 50 * rtol
        accepted.append(ctl.solver.step_controller.accepted)
    assert accepted[0] < accepted[1]


def test_adaptive_rejection_and_tick_schedule():
    cfg = _uniform_cfg("dp5", 1e-6, dt=5.0, total=5.0)
    ctl = _ctl(cfg)
    ctl.state.psi = np.ones((16, 16), dtype=complex)
    ctl.run()
    assert ctl.solver.step_controller.rejected > 0
    assert ctl.state.step == 1 and ctl.state.t == ctl.state.t0 + cfg.solver.dt
    assert np.max(np.abs(ctl.state.psi - _exact_uniform(cfg))) <= 50 * cfg.solver.rtol


@pytest.mark.parametrize("chunk", [1, 4])
def test_adaptive_trials_match_oracle_sequence(chunk):
    """Each advance of ``chunk`` ticks lands on its own target (abstract_solver.py:86-96)."""
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__tableau="dp5", grid__nx=32, grid__ny=32,
                            grid__lx=30.0, grid__ly=30.0, solver__dt=0.5, solver__total_time=2.0,
                            solver__adaptive=True, solver__rtol=1e-6, solver__atol=1e-9, laser__P0=0.3,
                            laser__sigma_space=6.0, noise__stochastic="none")
    ctl = _ctl(cfg)
    for _ in range(4 // chunk):
        ctl.solver.advance(ctl.model, ctl.state, chunk)
    assert ctl.state.step == 4 and ctl.state.t == ctl.state.t0 + 2.0
    g = O.Grid(32, 32, 30.0, 30.0)
    p = oracle_physics(cfg)
    psi = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed).reshape(1, 1, 32, 32)
    res = (np.zeros((1, 1, 32, 32)),)
    env = O.gaussian_profile(g, 6.0)
    prob = O.LawsonProblem("single", p, g, tableau="dp5", adaptive=True, rtol=1e-6, atol=1e-9)
    psi, res, ctl_o = O.lawson_advance_adaptive(prob, psi, res, lambda t: env * 0.3, 0.0, 0.5, 4, chunk=chunk)
    sc = ctl.solver.step_controller
    assert (sc.accepted, sc.rejected) == (ctl_o.accepted, ctl_o.rejected)
    assert O.rel_l2(ctl.state.field.get(), psi) <= 1e-10
