"""split-step-fft-hip (CPU order) and split-step-fft-kouter-hip against the oracle."""
import numpy as np
import pytest

from oracle import gpe_oracle as O
from tests.helpers import oracle_physics, quickstart_config

pytestmark = pytest.mark.gpu


def _oracle_split(cfg, n_steps, kouter=False, noise=False):
    g = O.Grid(cfg.grid.nx, cfg.grid.ny, cfg.grid.lx, cfg.grid.ly)
    prec = cfg.solver.precision
    real = np.float32 if prec == "single" else np.float64
    kind = cfg.reservoir.reservoir_type
    p = oracle_physics(cfg)
    psi = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed, prec).reshape(1, 1, *g.shape)
    nf = {"none": 0, "single": 1, "double": 2, "quadratic-double": 2}[kind]
    res = tuple(np.zeros((1, 1, *g.shape), dtype=real) for _ in range(nf))
    env = O.gaussian_profile(g, cfg.laser.sigma_space).astype(real)
    pump_at = lambda t: env * cfg.laser.P0
    nz = None
    if noise:
        nz = dict(seeds=[cfg.noise.seed], gamma=[p.gamma_C], scattering=[p.R],
                  strength=[cfg.noise.stochastic_strength], volume=g.dx * g.dy, precision=prec)
    dt = cfg.solver.dt
    fn = O.split_step_kouter if kouter else O.split_step_cpu
    for s in range(n_steps):
        psi, res = fn(kind, p, g, psi, res, pump_at, s * dt, dt, s, nz)
    return psi, res


def _cfg(**kw):
    base = dict(solver__method="split-step-fft-hip", grid__nx=128, grid__ny=96, grid__lx=64.0, grid__ly=48.0,
                laser__P0=0.2, laser__sigma_space=10.0, solver__dt=5e-3, solver__total_time=0.25)
    base.update(kw)
    return quickstart_config(**base)


def _run(cfg):
    from polarism_amd import SimulationController
    ctl = SimulationController(cfg)
    ctl.run()
    return ctl


@pytest.mark.parametrize("method,kouter", [("split-step-fft-hip", False), ("split-step-fft-kouter-hip", True)])
@pytest.mark.parametrize("kind", ["single", "double", "quadratic-double", "none"])
def test_split_step_fp64_matches_oracle(method, kouter, kind):
    cfg = _cfg(solver__method=method, reservoir__reservoir_type=kind)
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_split(cfg, 50, kouter)
    e = O.rel_l2(ctl.state.field.get(), psi_ref)
    assert e <= 1e-10, e
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= 1e-10


@pytest.mark.parametrize("method,kouter", [("split-step-fft-hip", False), ("split-step-fft-kouter-hip", True)])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_split_step_with_wigner_noise_matches_oracle(method, kouter, precision, tol):
    cfg = _cfg(solver__method=method, solver__precision=precision, noise__stochastic="wigner")
    cfg.noise.seed = 1234
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_split(cfg, 50, kouter, noise=True)
    psi = ctl.state.field.get()
    assert O.rel_l2(psi, psi_ref) <= tol
    assert O.rel_l2(np.abs(psi) ** 2, np.abs(psi_ref) ** 2) <= tol
    assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res_ref[0]) <= tol


def test_orderings_agree_to_second_order():
    """The two orderings differ at O(dt^2) (reference test_split_step_fft_cuda.py:190-224)."""
    errs = []
    for dt in (0.02, 0.01):
        # a strong field so that the reservoir depletion and the nonlinearity do not commute with K
        kw = dict(solver__dt=dt, solver__total_time=2.0, laser__P0=0.5, noise__initial_amplitude=0.5,
                  physics__g_C=0.05)
        a = _run(_cfg(**kw))
        b = _run(_cfg(solver__method="split-step-fft-kouter-hip", **kw))
        errs.append(O.rel_l2(a.state.field.get(), b.state.field.get()))
    assert errs[1] < errs[0] / 2.5


def test_split_step_graph_equals_direct_launches_bitwise():
    a = _run(_cfg(noise__stochastic="wigner"))
    b = _run(_cfg(noise__stochastic="wigner", solver__cuda_graphs=False))
    np.testing.assert_array_equal(a.state.field.get(), b.state.field.get())


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("pulsed", [False, True])
@pytest.mark.parametrize("graphs", [True, False])
def test_chained_steps_equal_separate_steps_bitwise(precision, pulsed, graphs):
    """advance(n) chains the CPU-order steps (one row pass between consecutive steps instead of
    two, the field kept in the scratch buffer); n calls of step() run each step whole.  The
    chain performs the same operations in the same order, so field and reservoir agree bit for
    bit, with noise and with a time-dependent (pulsed) pump whose coefficients differ per
    half-step."""
    from polarism_amd import SimulationController
    kw = dict(noise__stochastic="wigner", solver__precision=precision, grid__nx=64, grid__ny=48, grid__lx=32.0,
              grid__ly=24.0, solver__cuda_graphs=graphs)
    if pulsed:
        kw.update(laser__laser_type="pulse-gaussian")
    runs = []
    for chained in (True, False):
        cfg = _cfg(**kw)
        if pulsed:
            cfg.laser.sigma_time = 0.02
            cfg.laser.P0 = 0.4
        ctl = SimulationController(cfg)
        if chained:
            ctl.solver.advance(ctl.model, ctl.state, 24)
        else:
            for _ in range(24):
                ctl.solver.step(ctl.model, ctl.state, cfg.solver.dt)
        runs.append((ctl.state.field.get(), [r.get() for r in ctl.reservoir.get_state()], ctl.state.step))
    (fa, ra, sa), (fb, rb, sb) = runs
    assert sa == sb == 24
    np.testing.assert_array_equal(fa, fb)
    for x, y in zip(ra, rb):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_split_step_with_noise_config_window_matches_oracle(precision, tol):
    """The split-step-with-noise configuration at full size (512^2, lx = ly = 200 um, Gaussian
    sigma 20 um, P0 = 2 P_th, Wigner noise seed 1234, dt 5e-3 ps) over a 60-step window.

    psi, density, norm and reservoir are held to the reference's CPU-GPU bar (1e-10 FP64,
    1e-5 FP32, tests/crosscheck/test_cpu_gpu_agreement.py:12-13).  The 240 FP32 transforms of
    this window leave ~2e-6 on psi and density and ~3e-7 on the norm; the norm holds because
    the FP32 twiddles carry tails (without them each 2-D transform lost ~8e-8 of the energy,
    1.1e-5 over the window)."""
    cfg = _cfg(grid__nx=512, grid__ny=512, grid__lx=200.0, grid__ly=200.0, laser__sigma_space=20.0,
               laser__P0=2 * 0.02075, solver__precision=precision, solver__total_time=0.3,
               noise__stochastic="wigner", noise__seed=1234)
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_split(cfg, 60, noise=True)
    e = O.rel_l2(ctl.state.field.get(), psi_ref)
    d = O.rel_l2(np.abs(ctl.state.field.get()) ** 2, np.abs(psi_ref) ** 2)
    norm, norm_ref = float((np.abs(ctl.state.field.get()) ** 2).sum()), float((np.abs(psi_ref) ** 2).sum())
    en = abs(norm - norm_ref) / norm_ref
    print(f"512^2 split-step + noise {precision}: psi {e:.3e} density {d:.3e} norm {en:.3e}")
    assert e <= tol and d <= tol and en <= tol
    assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res_ref[0]) <= tol


def _oracle_split_spinor(cfg, n_steps, kouter=False, noise=False):
    g = O.Grid(cfg.grid.nx, cfg.grid.ny, cfg.grid.lx, cfg.grid.ly)
    prec = cfg.solver.precision
    real = np.float32 if prec == "single" else np.float64
    kind = cfg.reservoir.reservoir_type
    p = oracle_physics(cfg)
    plus = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed, prec)
    minus = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed + 1, prec)
    psi = np.stack((plus, minus))[None]
    C_res = cfg.reservoir.components or 2
    nf = {"none": 0, "single": 1, "double": 2, "quadratic-double": 2}[kind]
    res = tuple(np.zeros((1, C_res, *g.shape), dtype=real) for _ in range(nf))
    env = O.gaussian_profile(g, cfg.laser.sigma_space).astype(real)
    pump_at = lambda t: env * cfg.laser.P0
    cp = cfg.laser.circular_polarization
    w = np.asarray([(1 + cp) / 2, (1 - cp) / 2], dtype=real).reshape(1, 2, 1, 1)
    nz = None
    if noise:
        nz = dict(seeds=[cfg.noise.seed], gamma=[p.gamma_C], scattering=[p.R],
                  strength=[cfg.noise.stochastic_strength], volume=g.dx * g.dy, precision=prec)
    dt = cfg.solver.dt
    for s in range(n_steps):
        psi, res = O.split_step_spinor(kind, p, g, psi, res, pump_at, s * dt, dt, s, w, C_res, nz, kouter=kouter)
    return psi, res


def _spinor_cfg(**kw):
    base = dict(physics__spin="spinor", physics__te_tm=0.05, physics__zeeman=0.02, physics__alpha_2=-1e-4,
                laser__circular_polarization=0.3)
    base.update(kw)
    return _cfg(**base)


@pytest.mark.parametrize("method,kouter", [("split-step-fft-hip", False), ("split-step-fft-kouter-hip", True)])
@pytest.mark.parametrize("kind,res_components", [("single", 2), ("double", 2), ("quadratic-double", 2),
                                                 ("single", 1), ("double", 1), ("none", 2)])
def test_spinor_split_step_fp64_matches_oracle(method, kouter, kind, res_components):
    """psi+ / psi- with TE-TM, Zeeman, cross-spin interaction and an elliptic pump on 128 x 96:
    50 steps against the oracle, rel-L2 <= 1e-10 on both components and every reservoir field,
    per-component and shared reservoirs."""
    cfg = _spinor_cfg(solver__method=method, reservoir__reservoir_type=kind, reservoir__components=res_components)
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_split_spinor(cfg, 50, kouter)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= 1e-10
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= 1e-10


@pytest.mark.parametrize("ny,precision,tol", [(2048, "double", 1e-10), (2048, "single", 1e-5),
                                               (4096, "single", 1e-5)])
def test_tall_spinor_columns_match_oracle(ny, precision, tol):
    """Columns of 2048 and 4096 points: the spinor column groups of 256 threads x 16 and x 32
    elements (one column of psi+ and one of psi- per group); 10 steps against the oracle."""
    cfg = _spinor_cfg(grid__nx=32, grid__ny=ny, grid__lx=16.0, grid__ly=0.5 * ny, solver__precision=precision,
                      solver__total_time=0.05)
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_split_spinor(cfg, 10)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= tol
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= tol


@pytest.mark.parametrize("method,kouter", [("split-step-fft-hip", False), ("split-step-fft-kouter-hip", True)])
@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_spinor_split_step_with_wigner_noise_matches_oracle(method, kouter, precision, tol):
    """Wigner noise on both components (streams 1 and 2) with per-component reservoirs."""
    cfg = _spinor_cfg(solver__method=method, solver__precision=precision, noise__stochastic="wigner")
    cfg.noise.seed = 1234
    ctl = _run(cfg)
    psi_ref, res_ref = _oracle_split_spinor(cfg, 50, kouter, noise=True)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= tol
    for d, r in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(d.get(), r) <= tol
