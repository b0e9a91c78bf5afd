"""rk4-fdm-hip against the oracle restatement of the reference rk4-fdm."""
import numpy as np
import pytest

from oracle import gpe_oracle as O
from tests.helpers import oracle_rk4_run, quickstart_config

pytestmark = pytest.mark.gpu


def _run(cfg, members=None):
    from polarism_amd import SimulationController
    ctl = SimulationController(cfg, members=members)
    ctl.run()
    return ctl


def test_quickstart_fp64_matches_oracle():
    """Config 1: 128^2, 2000 steps, complex128; rel-L2 <= 1e-10 on psi and nR."""
    cfg = quickstart_config()
    ctl = _run(cfg)
    psi_ref, res_ref = oracle_rk4_run(cfg, 2000)
    psi = ctl.state.field.get()
    nR = ctl.reservoir.get_state()[0].get()
    e_psi = O.rel_l2(psi, psi_ref)
    e_res = O.rel_l2(nR, res_ref[0])
    print(f"quickstart fp64 rel-L2 psi={e_psi:.3e} nR={e_res:.3e}")
    assert e_psi <= 1e-10 and e_res <= 1e-10


def test_quickstart_fp32_matches_oracle_and_reports_fp64_error():
    cfg = quickstart_config(solver__precision="single", solver__total_time=0.5)
    ctl = _run(cfg)
    psi_ref, res_ref = oracle_rk4_run(cfg, 500)
    psi = ctl.state.field.get()
    nR = ctl.reservoir.get_state()[0].get()
    e_psi = O.rel_l2(psi, psi_ref)
    dens = O.rel_l2(np.abs(psi) ** 2, np.abs(psi_ref) ** 2)
    cfg64 = quickstart_config(solver__total_time=0.5)
    psi64, _ = oracle_rk4_run(cfg64, 500)
    e64 = O.rel_l2(psi.astype(np.complex128), psi64)
    print(f"fp32 rel-L2 psi={e_psi:.3e} density={dens:.3e} nR={O.rel_l2(nR, res_ref[0]):.3e} vs-fp64={e64:.3e}")
    assert e_psi <= 1e-5 and dens <= 1e-5 and O.rel_l2(nR, res_ref[0]) <= 1e-5


@pytest.mark.parametrize("kind", ["double", "quadratic-double", "none"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_reservoir_kinds_match_oracle(kind, precision, tol):
    cfg = quickstart_config(reservoir__reservoir_type=kind, solver__precision=precision, solver__total_time=0.2,
                            grid__nx=96, grid__ny=80, laser__P0=0.2)
    cfg.solver.dt = 2e-3
    ctl = _run(cfg)
    psi_ref, res_ref = oracle_rk4_run(cfg, 100)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= tol
    for dev, ref in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(dev.get(), ref) <= tol


def test_pulsed_pump_and_eta_match_oracle():
    cfg = quickstart_config(laser__laser_type="pulse-gaussian", solver__total_time=0.2, grid__nx=64, grid__ny=48,
                            physics__kinetic_relaxation_eta=0.04)
    cfg.laser.sigma_time = 0.03
    cfg.laser.P0 = 0.4
    cfg.solver.dt = 2e-3
    ctl = _run(cfg)
    psi_ref, res_ref = oracle_rk4_run(cfg, 100)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= 1e-10
    assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res_ref[0]) <= 1e-10


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_wigner_noise_run_matches_oracle(precision, tol):
    cfg = quickstart_config(noise__stochastic="wigner", solver__precision=precision, solver__total_time=0.05,
                            grid__nx=64, grid__ny=64, laser__P0=0.05)
    cfg.noise.seed = 1234
    cfg.solver.dt = 5e-3
    ctl = _run(cfg)
    psi_ref, res_ref = oracle_rk4_run(cfg, 10, noise=True)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= tol
    assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res_ref[0]) <= tol


def test_graph_chunks_equal_single_launches_bitwise():
    cfg = quickstart_config(solver__total_time=0.05, noise__stochastic="wigner", grid__nx=64, grid__ny=64)
    a = _run(cfg)
    cfg2 = quickstart_config(solver__total_time=0.05, noise__stochastic="wigner", grid__nx=64, grid__ny=64,
                             solver__cuda_graphs=False)
    b = _run(cfg2)
    np.testing.assert_array_equal(a.state.field.get(), b.state.field.get())
    np.testing.assert_array_equal(a.reservoir.get_state()[0].get(), b.reservoir.get_state()[0].get())


def test_batch_member_equals_single_run_bitwise():
    cfg = quickstart_config(solver__total_time=0.05, grid__nx=64, grid__ny=64, noise__stochastic="wigner")
    members = [{"laser.P0": 0.02, "noise.seed": 5}, {"laser.P0": 0.05, "noise.seed": 6},
               {"laser.P0": 0.03, "physics.g_C": 0.002, "noise.seed": 7}]
    batch = _run(cfg, members)
    for m, over in enumerate(members):
        single = _run(cfg, [over])
        np.testing.assert_array_equal(batch.state.field.get()[m], single.state.field.get()[0])


@pytest.mark.parametrize("precision", ["double", "single"])
def test_tuned_form_equals_lds_tiles(precision):
    """The controller's whole-step form chosen by the autotuner (strips or LDS tiles) gives the
    LDS tiles' result bit for bit on a periodic grid, Wigner noise included."""
    from polarism_amd import SimulationController
    cfg = quickstart_config(solver__total_time=0.03, grid__nx=96, grid__ny=72, solver__precision=precision,
                            noise__stochastic="wigner")
    runs = []
    for tuned in (True, False):
        ctl = SimulationController(cfg)
        ctl.solver.autotune = tuned
        ctl.run()
        runs.append((ctl.state.field.get(), ctl.reservoir.get_state()[0].get(), ctl.solver.engine.strip_h))
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    np.testing.assert_array_equal(runs[0][1], runs[1][1])
    assert runs[1][2] is None


@pytest.mark.parametrize("tile_h", [8, 32])
def test_tile_height_does_not_change_results(tile_h):
    from polarism_amd import SimulationController
    cfg = quickstart_config(solver__total_time=0.02, grid__nx=96, grid__ny=72)
    ref = _run(cfg)
    ctl = SimulationController(cfg)
    ctl.solver.tile_h = tile_h
    ctl.run()
    np.testing.assert_array_equal(ctl.state.field.get(), ref.state.field.get())


def _oracle_spinor_rk4(cfg, n_steps, res_components=2):
    g = O.Grid(cfg.grid.nx, cfg.grid.ny, cfg.grid.lx, cfg.grid.ly, boundary=cfg.grid.grid_type)
    prec = cfg.solver.precision
    real = np.float32 if prec == "single" else np.float64
    kind = cfg.reservoir.reservoir_type
    from tests.helpers import oracle_physics
    p = oracle_physics(cfg)
    a = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed, prec)
    b = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed + 1, prec)
    psi = np.stack((a, b))[None]
    nf = {"none": 0, "single": 1, "double": 2, "quadratic-double": 2}[kind]
    res = tuple(np.zeros((1, res_components, *g.shape), dtype=real) for _ in range(nf))
    env = O.gaussian_profile(g, cfg.laser.sigma_space, cfg.laser.x0, cfg.laser.y0).astype(real)
    cp = cfg.laser.circular_polarization
    weights = np.asarray([(1 + cp) / 2, (1 - cp) / 2], dtype=real).reshape(1, 2, 1, 1)
    dt = cfg.solver.dt
    for s in range(n_steps):
        psi, res = O.rk4_fdm_step(kind, p, g, psi, res, lambda t: env * cfg.laser.P0, s * dt, dt,
                                  pump_weights=weights, reservoir_components=res_components)
    return psi, res


@pytest.mark.parametrize("kind", ["single", "double"])
@pytest.mark.parametrize("res_components", [2, 1])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_spinor_rk4_matches_oracle(kind, res_components, precision, tol):
    cfg = quickstart_config(grid__nx=96, grid__ny=64, grid__lx=48.0, grid__ly=32.0, solver__dt=2e-3,
                            solver__total_time=0.2, physics__spin="spinor", physics__te_tm=0.05,
                            physics__zeeman=0.02, physics__alpha_2=-1e-4, laser__circular_polarization=0.3,
                            laser__P0=0.3, laser__sigma_space=8.0, reservoir__reservoir_type=kind,
                            solver__precision=precision, noise__initial_amplitude=0.05)
    cfg.reservoir.components = res_components
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_spinor_rk4(cfg, 100, res_components)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= tol
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= tol


@pytest.mark.parametrize("res_components", [2, 1])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_spinor_closed_interval_matches_oracle(res_components, precision, tol):
    """Spinor on the closed interval: the halo mirrored about the wall nodes feeds the Laplacian
    (model.py:227-248) and the reflect-padded TE-TM differences (model.py:196-205); the pump sits
    off-centre and TE-TM is strong so the walls matter."""
    cfg = quickstart_config(grid__nx=80, grid__ny=56, grid__lx=40.0, grid__ly=28.0,
                            grid__grid_type="closed-interval", solver__dt=2e-3, solver__total_time=0.2,
                            physics__spin="spinor", physics__te_tm=0.2, physics__zeeman=0.02,
                            physics__alpha_2=-1e-4, laser__circular_polarization=0.3, laser__P0=0.3,
                            laser__sigma_space=6.0, laser__x0=14.0, laser__y0=-9.0,
                            reservoir__reservoir_type="single", solver__precision=precision,
                            noise__initial_amplitude=0.05)
    cfg.reservoir.components = res_components
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_spinor_rk4(cfg, 100, res_components)
    e = O.rel_l2(ctl.state.field.get(), psi_ref)
    print(f"spinor closed-interval {precision} res_components={res_components}: rel-L2 psi={e:.3e}")
    assert e <= tol
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= tol


def test_spinor_without_coupling_equals_two_scalar_runs():
    """Zeeman, TE-TM and alpha_2 off, shared nothing: each spin evolves as the scalar field
    (reference tests/unit/test_spin_8a.py:78-103)."""
    kw = dict(grid__nx=64, grid__ny=64, solver__total_time=0.05, laser__P0=0.03, noise__initial_amplitude=0.05)
    spin = quickstart_config(physics__spin="spinor", **kw)
    s = _run(spin)
    # circular_polarization 0 gives each spin its own reservoir and half the pump
    kw["laser__P0"] = 0.015
    half = quickstart_config(**kw)
    for c, seed in enumerate((0, 1)):
        half.noise.seed = seed
        ref = _run(half)
        np.testing.assert_allclose(s.state.field.get()[0, c], ref.state.field.get()[0, 0], rtol=1e-12, atol=1e-18)


@pytest.mark.parametrize("grid_type", ["periodic", "closed-interval"])
@pytest.mark.parametrize("laplacian", ["five-point", "isotropic-9pt"])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_boundaries_and_stencils_match_oracle(grid_type, laplacian, precision, tol):
    """Closed-interval mirrored edges (model.py:227-248) and the isotropic 9-point stencil
    (rk4_cuda_solver.py:719-770) against the oracle, with the pump off-centre so the edges matter."""
    cfg = quickstart_config(grid__nx=72, grid__ny=72, grid__lx=36.0, grid__ly=36.0, grid__grid_type=grid_type,
                            solver__laplacian=laplacian, solver__precision=precision, solver__total_time=0.3,
                            laser__P0=0.3, laser__sigma_space=6.0, laser__x0=12.0, laser__y0=-9.0,
                            noise__initial_amplitude=0.05)
    cfg.solver.dt = 2e-3
    ctl = _run(cfg)
    psi_ref, res_ref = oracle_rk4_run(cfg, 150)
    e = O.rel_l2(ctl.state.field.get(), psi_ref)
    print(f"{grid_type} {laplacian} {precision}: rel-L2 psi={e:.3e}")
    assert e <= tol
    assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res_ref[0]) <= tol


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_complex_potential_matches_oracle(precision, tol):
    """A static complex potential V enters (V + gC|psi|^2 + ...) psi (model.py:89-127)."""
    from polarism_amd import SimulationController
    cfg = quickstart_config(grid__nx=64, grid__ny=64, grid__lx=40.0, grid__ly=40.0, solver__precision=precision,
                            solver__total_time=0.2, laser__P0=0.2, noise__initial_amplitude=0.05)
    cfg.solver.dt = 2e-3
    g = O.Grid(64, 64, 40.0, 40.0)
    pot = (0.3 * np.exp(-(g.X ** 2 + g.Y ** 2) / 50.0) - 0.02j * (g.X / 20.0) ** 2).astype(np.complex128)
    ctl = SimulationController(cfg)
    ctl.model.potential = pot
    ctl.run()
    real = np.float32 if precision == "single" else np.float64
    cplx = np.complex64 if precision == "single" else np.complex128
    psi_ref, res_ref = oracle_rk4_run(cfg, 100, potential=pot.astype(cplx))
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= tol
    assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res_ref[0].astype(real)) <= tol
