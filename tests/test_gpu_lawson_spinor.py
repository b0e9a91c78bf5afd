"""ifrk-fft-hip on the spinor Lawson DP5 workload (1024^2, lx = ly = 300 um, TE-TM 0.05, Zeeman 0.02,
alpha_2 = -0.1 g_C, elliptic ring pump r 15 um / width 3 um, double reservoir with one reservoir per
component), adaptive trials with rejections, Wigner noise in the fused fixed step, and the embedded
error norm of a two-member batch against its closed form.

Every HIP result is compared with the oracle on the same inputs:
complex128 at 1e-10 rel-L2, complex64 at 1e-5 on density and norm, and the same accepted /
rejected trial counts.  The complex64 runs also print their error against a complex128 oracle."""
import numpy as np
import pytest

from oracle import gpe_oracle as O
from tests.helpers import oracle_physics, quickstart_config

pytestmark = pytest.mark.gpu


def _ctl(cfg, members=None):
    from polarism_amd import SimulationController
    return SimulationController(cfg, members=members)


def spinor_ring_cfg(n=1024, lx=300.0, precision="single", rtol=1e-5, dt=1.0, ticks=3, seed=3):
    cfg = quickstart_config(grid__nx=n, grid__ny=n, grid__lx=lx, grid__ly=lx, solver__method="ifrk-fft-hip",
                            solver__tableau="dp5", solver__adaptive=True, solver__rtol=rtol, solver__atol=1e-8,
                            solver__dt=dt, solver__total_time=dt * ticks, solver__precision=precision,
                            physics__spin="spinor", physics__te_tm=0.05, physics__zeeman=0.02,
                            reservoir__reservoir_type="double", laser__laser_type="profile", laser__P0=0.1,
                            laser__circular_polarization=0.5, noise__seed=seed)
    cfg.physics.alpha_2 = -0.1 * cfg.physics.g_C
    cfg.laser.profile = {"type": "ring", "radius": 15.0, "width": 3.0}
    return cfg


def _oracle_spinor_adaptive(cfg, ticks, precision=None, rtol=None):
    prec = precision or cfg.solver.precision
    real = np.float32 if prec == "single" else np.float64
    g = O.Grid(cfg.grid.nx, cfg.grid.ny, cfg.grid.lx, cfg.grid.ly)
    p = oracle_physics(cfg)
    seed = cfg.noise.seed
    psi = np.stack((O.initial_psi(g.shape, cfg.noise.initial_amplitude, seed, prec),
                    O.initial_psi(g.shape, cfg.noise.initial_amplitude, seed + 1, prec)))[None]
    res = tuple(np.zeros((1, 2) + g.shape, dtype=real) for _ in range(2))
    prof = cfg.laser.profile
    env = O.ring_profile(g, prof["radius"], prof["width"]).astype(real)
    cp = cfg.laser.circular_polarization
    weights = np.asarray([(1 + cp) / 2, (1 - cp) / 2], dtype=real).reshape(1, 2, 1, 1)
    prob = O.LawsonProblem("double", p, g, precision=prec, tableau=cfg.solver.tableau, adaptive=True,
                           rtol=rtol or cfg.solver.rtol, atol=cfg.solver.atol, pump_weights=weights,
                           reservoir_components=2)
    P0 = cfg.laser.P0
    return O.lawson_advance_adaptive(prob, psi, res, lambda t: env * P0, 0.0, cfg.solver.dt, ticks)


def _density_norm_errors(psi, ref):
    d, dr = np.abs(psi) ** 2, np.abs(ref) ** 2
    return O.rel_l2(d, dr), abs(d.sum() - dr.sum()) / dr.sum()


def _run_ticks(ctl, ticks):
    for _ in range(ticks):
        ctl.solver.advance(ctl.model, ctl.state, 1)
    return ctl.solver.step_controller


def test_spinor_ring_dp5_complex64_1024_matches_oracle():
    """Three 1 ps ticks of the complex64 workload (rtol 1e-5, the smallest the reference accepts in
    single precision): density and norm within 1e-5 of the complex64 oracle, same trial counts."""
    cfg = spinor_ring_cfg()
    ctl = _ctl(cfg)
    sc = _run_ticks(ctl, 3)
    assert ctl.state.step == 3 and ctl.state.t == 3.0
    psi_ref, res_ref, octl = _oracle_spinor_adaptive(cfg, 3)
    assert (sc.accepted, sc.rejected) == (octl.accepted, octl.rejected)
    psi = ctl.state.field.get()
    e_d, e_n = _density_norm_errors(psi, psi_ref)
    e_psi = O.rel_l2(psi, psi_ref)
    e_res = [O.rel_l2(a.get(), b) for a, b in zip(ctl.reservoir.get_state(), res_ref)]
    psi64, _, _ = _oracle_spinor_adaptive(cfg, 3, precision="double")
    f_d, f_n = _density_norm_errors(psi, psi64)
    print(f"c64 1024^2 spinor DP5: trials {sc.accepted}+{sc.rejected}; vs c64 oracle: density {e_d:.2e} "
          f"norm {e_n:.2e} psi {e_psi:.2e} res {max(e_res):.2e}; vs FP64 oracle: density {f_d:.2e} norm {f_n:.2e}",
          flush=True)
    assert e_d <= 1e-5 and e_n <= 1e-5
    assert max(e_res) <= 1e-5


def test_spinor_ring_dp5_complex128_1024_matches_oracle():
    """One tick in complex128 at rtol 1e-6: the same trials and 1e-10 on psi and the reservoirs."""
    cfg = spinor_ring_cfg(precision="double", rtol=1e-6)
    ctl = _ctl(cfg)
    sc = _run_ticks(ctl, 1)
    psi_ref, res_ref, octl = _oracle_spinor_adaptive(cfg, 1)
    assert (sc.accepted, sc.rejected) == (octl.accepted, octl.rejected)
    assert O.rel_l2(ctl.state.field.get(), psi_ref) <= 1e-10
    for a, b in zip(ctl.reservoir.get_state(), res_ref):
        assert O.rel_l2(a.get(), b) <= 1e-10


@pytest.mark.parametrize("precision", ["single", "double"])
def test_spinor_adaptive_double_reservoir_with_rejections(precision):
    """10 ps ticks on a 64^2 box force rejected trials; the accept / reject sequence and the
    state follow the oracle in both precisions."""
    cfg = spinor_ring_cfg(n=64, lx=80.0, precision=precision, rtol=1e-5, dt=10.0, ticks=3)
    ctl = _ctl(cfg)
    sc = _run_ticks(ctl, 3)
    psi_ref, res_ref, octl = _oracle_spinor_adaptive(cfg, 3)
    assert octl.rejected > 0
    assert (sc.accepted, sc.rejected) == (octl.accepted, octl.rejected)
    psi = ctl.state.field.get()
    if precision == "double":
        assert O.rel_l2(psi, psi_ref) <= 1e-10
        for a, b in zip(ctl.reservoir.get_state(), res_ref):
            assert O.rel_l2(a.get(), b) <= 1e-10
    else:
        e_d, e_n = _density_norm_errors(psi, psi_ref)
        psi64, _, _ = _oracle_spinor_adaptive(cfg, 3, precision="double")
        f_d, _ = _density_norm_errors(psi, psi64)
        print(f"c64 64^2 spinor DP5 with rejections: density {e_d:.2e} norm {e_n:.2e}; vs FP64 {f_d:.2e}", flush=True)
        assert e_d <= 1e-5 and e_n <= 1e-5


# ---------------------------------------------------------------- Wigner noise in the fused fixed step
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("spinor", [False, True])
def test_lawson_fixed_step_wigner_noise_matches_oracle(precision, spinor):
    """ifrk_cuda_fused.py:908-923: after each fixed step, inverse transform, the Wigner increment of
    that step index with the final reservoir, forward transform."""
    kw = dict(physics__spin="spinor", physics__te_tm=0.05, physics__zeeman=0.02, physics__alpha_2=-1e-4,
              laser__circular_polarization=0.3) if spinor else {}
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__tableau="dp5", grid__nx=64, grid__ny=48,
                            grid__lx=40.0, grid__ly=30.0, solver__dt=0.02, solver__total_time=0.2,
                            solver__precision=precision, laser__P0=0.3, laser__sigma_space=8.0,
                            reservoir__reservoir_type="single", noise__stochastic="wigner", noise__seed=1234, **kw)
    ctl = _ctl(cfg)
    ctl.run()
    assert ctl.state.step == 10
    prec = precision
    real = np.float32 if prec == "single" else np.float64
    g = O.Grid(64, 48, 40.0, 30.0)
    p = oracle_physics(cfg)
    C = 2 if spinor else 1
    psi = O.initial_psi(g.shape, cfg.noise.initial_amplitude, 1234, prec)
    psi = (np.stack((psi, O.initial_psi(g.shape, cfg.noise.initial_amplitude, 1235, prec)))[None] if spinor
           else psi.reshape(1, 1, *g.shape))
    res = (np.zeros((1, C) + g.shape, dtype=real),)
    env = O.gaussian_profile(g, 8.0).astype(real)
    weights = None
    if spinor:
        weights = np.asarray([0.65, 0.35], dtype=real).reshape(1, 2, 1, 1)
    prob = O.LawsonProblem("single", p, g, precision=prec, tableau="dp5", pump_weights=weights,
                           reservoir_components=C)
    noise = dict(seeds=[1234], gamma=[p.gamma_C], scattering=[p.R], strength=[1.0], volume=g.dx * g.dy,
                 precision=prec)
    for s in range(10):
        _, _, psi, res = prob.step(psi, res, lambda t: env * 0.3, s * 0.02, 0.02, noise=noise, step=s)
    out = ctl.state.field.get()
    if precision == "double":
        assert O.rel_l2(out, psi) <= 1e-10
        assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res[0]) <= 1e-10
    else:
        e_d, e_n = _density_norm_errors(out, psi)
        print(f"c64 Lawson DP5 + Wigner (spinor={spinor}): psi {O.rel_l2(out, psi):.2e} density {e_d:.2e} "
              f"norm {e_n:.2e}", flush=True)
        assert O.rel_l2(out, psi) <= 1e-5 and e_d <= 1e-5 and e_n <= 1e-5


# ---------------------------------------------------------------- embedded error norm, two members
@pytest.mark.parametrize("tableau", ["dp5", "dop853"])
def test_error_norm_of_two_member_batch_matches_closed_form(tableau):
    """The HIP trial's reported error on the linear uniform problem of the oracle pin: the largest
    member's norm from the embedded stability functions (psi and reservoir points both counted)."""
    from tests.test_oracle_pins import _increment, _stability
    n = 16
    h = 0.8 if tableau == "dp5" else 2.0
    V, P0, rtol, atol = 0.05, 0.02, 1e-6, 1e-3
    cfg = quickstart_config(solver__method="ifrk-fft-hip", solver__tableau=tableau, grid__nx=n, grid__ny=n,
                            grid__lx=16.0, grid__ly=16.0, solver__dt=h, solver__total_time=h, solver__adaptive=True,
                            solver__rtol=rtol, solver__atol=atol, physics__g_C=0.0, physics__R=0.0,
                            physics__g_R=0.0, physics__gamma_C=0.6, physics__gamma_R=0.3,
                            laser__laser_type="profile", laser__P0=P0, reservoir__reservoir_type="single",
                            noise__stochastic="none")
    cfg.laser.profile = {"type": "uniform"}
    ctl = _ctl(cfg, members=[{}, {}])
    a = np.array([0.04 - 0.02j, 1.0 + 0.5j])
    n0 = np.array([0.01, 0.2])
    ctl.state.field = np.broadcast_to(a[:, None, None, None], (2, 1, n, n)).astype(complex)
    ctl.reservoir.set_state((np.broadcast_to(n0[:, None, None, None], (2, 1, n, n)).astype(float),))
    ctl.model.potential = np.full((n, n), V)
    report = ctl.solver.step(ctl.model, ctl.state, h)
    p = oracle_physics(cfg)
    tab = O.oracle_tableau(tableau)
    b = np.asarray(tab.b, dtype=float)
    z = h * (-1j * V / p.hbar - p.gamma_C / 2)
    zr = -p.gamma_R * h
    nstar = P0 / p.gamma_R
    psi1 = _stability(tab, b, z) * a
    res1 = nstar + _stability(tab, b, zr).real * (n0 - nstar)

    def sums(bhat):
        w = b - np.asarray(bhat, dtype=float)
        e_psi = np.abs(_increment(tab, w, z) * a) / (atol + rtol * np.maximum(np.abs(a), np.abs(psi1)))
        e_res = np.abs(_increment(tab, w, zr).real * (n0 - nstar)) / (atol + rtol * np.maximum(n0, res1))
        return n * n * (e_psi ** 2 + e_res ** 2)

    s5 = sums(tab.b_hat)
    member = np.sqrt(s5 / (2 * n * n)) if tab.b_hat3 is None else s5 / np.sqrt((s5 + 0.01 * sums(tab.b_hat3)) * 2 * n * n)
    assert member[1] > 1.5 * member[0]
    print(f"{tableau} two-member error: HIP {report.error:.12e} closed form {member.max():.12e}", flush=True)
    assert abs(report.error - member.max()) <= 1e-7 * member.max()
    assert report.accepted == (member.max() <= 1)


# ---------------------------------------------------------------- device-side trial control
@pytest.mark.parametrize("tableau", ["dp5", "dop853"])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_device_controlled_trials_equal_host_controlled_trials(tableau, precision):
    """The same adaptive run with the PI controller on the device (captured trial graphs) and on
    the host (a read-back per trial): the same accept / reject sequence, time and statistics, and
    states equal to rounding.  The two differ only in rounding: the controller's pow may differ in
    the last bit between libraries, and the host path's first stage reads the real-space field
    where the device path transforms it first.  In single precision the embedded error is a
    difference of FP32 stage values, so such rounding moves the error, and with it the next h,
    by ~1e-4 relative."""
    runs = []
    for device in (True, False):
        cfg = spinor_ring_cfg(n=64, lx=80.0, precision=precision, rtol=1e-5 if precision == "single" else 1e-7,
                              dt=10.0, ticks=3)
        cfg.solver.tableau = tableau
        ctl = _ctl(cfg)
        ctl.solver.device_trials = device
        sc = _run_ticks(ctl, 3)
        runs.append((ctl, dict(sc.statistics), sc.h_next, ctl.state.field.get(),
                     [r.get() for r in ctl.reservoir.get_state()]))
    (dev, s_dev, h_dev, psi_dev, res_dev), (host, s_host, h_host, psi_host, res_host) = runs
    print(f"{tableau} {precision}: device {s_dev} host {s_host}", flush=True)
    assert s_dev["rejected"] > 0
    assert (s_dev["accepted"], s_dev["rejected"]) == (s_host["accepted"], s_host["rejected"])
    assert dev.state.t == host.state.t == 30.0 and dev.state.step == host.state.step == 3
    assert abs(h_dev - h_host) <= (1e-9 if precision == "double" else 1e-3) * h_host
    tol = 1e-12 if precision == "double" else 1e-5
    assert O.rel_l2(psi_dev, psi_host) <= tol
    for a, b in zip(res_dev, res_host):
        assert O.rel_l2(a, b) <= tol
