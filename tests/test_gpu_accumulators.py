"""Device accumulators (momentum distribution, g2) against the same sums taken on the host."""
import numpy as np
import pytest

from tests.helpers import quickstart_config

pytestmark = pytest.mark.gpu


def _cfg(**kw):
    base = dict(solver__method="split-step-fft-hip", grid__nx=64, grid__ny=48, grid__lx=40.0, grid__ly=30.0,
                laser__P0=0.2, laser__sigma_space=8.0, solver__dt=5e-3, solver__total_time=0.05,
                noise__stochastic="wigner")
    base.update(kw)
    cfg = quickstart_config(**base)
    cfg.noise.seed = 77
    return cfg


@pytest.mark.parametrize("precision,spin", [("double", "scalar"), ("single", "scalar"), ("double", "spinor")])
def test_accumulated_sums_equal_host_sums_of_the_sampled_fields(precision, spin):
    """Samples every 2 steps from step 0 (six samples); the device FP64 sums, normalised and
    Wigner-corrected as the reference does, equal the host sums of the same fields."""
    from polarism_amd import SimulationController
    kw = dict(solver__precision=precision, physics__spin=spin)
    if spin == "spinor":
        kw.update(physics__te_tm=0.05, physics__zeeman=0.02)
    cfg = _cfg(**kw)
    cfg.result.accumulators = {"momentum-distribution": {"every": 0.01}, "g2": {"every": 0.01}}
    ctl = SimulationController(cfg)
    ctl.run()
    acc = ctl.accumulated

    ref = SimulationController(_cfg(**kw))
    fields = [ref.state.field.get().astype(np.complex128)]
    for _ in range(5):
        ref.solver.advance(ref.model, ref.state, 2)
        fields.append(ref.state.field.get().astype(np.complex128))
    shape = ref.state._public_shape
    g = ref.grid
    cell, cells, n = g.dx * g.dy, g.nx * g.ny, len(fields)
    strength = cfg.noise.stochastic_strength
    sum_k = sum(np.abs(np.fft.fft2(f, axes=(-2, -1))) ** 2 for f in fields).reshape(shape)
    nk = np.fft.fftshift(sum_k * (cell / (cells * n)), axes=(-2, -1)) - strength / 2
    rho = [np.abs(f.reshape(shape)) ** 2 for f in fields]
    delta = strength / cell
    mean = sum(rho) / n
    sq = sum(r * r for r in rho) / n
    sq, mean = sq - 2 * delta * mean + delta * delta / 2, mean - delta / 2
    tol = 1e-11 if precision == "double" else 1e-5
    np.testing.assert_allclose(acc["momentum_distribution"].value, nk, rtol=tol, atol=tol * np.max(np.abs(nk)))
    np.testing.assert_allclose(acc["density_mean"].value, mean, rtol=tol, atol=tol * np.max(np.abs(mean)))
    np.testing.assert_allclose(acc["density_squared_mean"].value, sq, rtol=tol, atol=tol * np.max(np.abs(sq)))
    assert acc["momentum_distribution"].wigner_corrected and acc["g2"].value.shape == shape
    assert acc["momentum_distribution"].dims[-2:] == ("ky", "kx")


def test_plane_wave_occupies_one_mode_with_unit_g2():
    """A plane wave on the grid with no gain, loss, interaction or pump stays a plane wave (RK4
    steps short enough that the amplitude error is ~1e-15): its whole particle number
    a^2 N dx dy sits in one mode, g1 has unit modulus and g2 = 1 everywhere."""
    from polarism_amd import SimulationController
    cfg = quickstart_config(solver__method="rk4-fdm-hip", grid__nx=32, grid__ny=32, grid__lx=32.0, grid__ly=32.0,
                            solver__dt=1e-5, solver__total_time=1e-4, laser__P0=0.0, reservoir__reservoir_type="none",
                            physics__gamma_C=0.0, physics__g_C=0.0)
    cfg.noise.enabled = False
    cfg.result.accumulators = {"momentum-distribution": {"every": 5e-5}, "g2": {}}
    ctl = SimulationController(cfg)
    g = ctl.grid
    a = 0.7
    kx, ky = 2 * np.pi * 3 / g.lx, -2 * np.pi * 2 / g.ly
    x = (np.arange(g.nx) - (g.nx - 1) / 2) * g.dx
    y = (np.arange(g.ny) - (g.ny - 1) / 2) * g.dy
    X, Y = np.meshgrid(x, y)
    ctl.state.psi = a * np.exp(1j * (kx * X + ky * Y))
    ctl.run()
    nk = ctl.accumulated["momentum_distribution"].value
    total = a * a * g.nx * g.ny * g.dx * g.dy
    peak = np.unravel_index(np.argmax(nk), nk.shape)
    assert peak == (g.ny // 2 - 2, g.nx // 2 + 3)
    np.testing.assert_allclose(nk[peak], total, rtol=1e-10)
    assert np.sum(nk) == pytest.approx(total, rel=1e-10)
    np.testing.assert_allclose(ctl.accumulated["g2"].value, 1.0, rtol=1e-10)
    np.testing.assert_allclose(np.abs(ctl.accumulated["g1"].value), 1.0, atol=1e-8)
