"""Checkpoints: a run resumed from its mid-run checkpoint ends bit for bit where the
uninterrupted run ends (simulation_controller.py:736-923)."""
import numpy as np
import pytest

from tests.helpers import quickstart_config

pytestmark = pytest.mark.gpu


def _pair(tmp_path, cfg_fn):
    from polarism_amd import SimulationController
    path = tmp_path / "ck.npz"
    a_cfg = cfg_fn()
    a_cfg.checkpoint.path = str(path)
    a_cfg.checkpoint.every = a_cfg.solver.total_time / 2
    a = SimulationController(a_cfg)
    a.run()
    # the file now holds the end state; write the mid-run one by re-running half the time
    half_cfg = cfg_fn()
    half_cfg.solver.total_time = a_cfg.solver.total_time / 2
    half_cfg.checkpoint.path = str(path)
    h = SimulationController(half_cfg)
    h.run()
    b_cfg = cfg_fn()
    b_cfg.checkpoint.resume_from = str(path)
    b_cfg.checkpoint.path = str(tmp_path / "ck_b.npz")
    b = SimulationController(b_cfg)
    assert b.state.step == h.state.step > 0
    b.run()
    return a, b


def _equal(a, b):
    np.testing.assert_array_equal(b.state.field.get(), a.state.field.get())
    for x, y in zip(a.reservoir.get_state(), b.reservoir.get_state()):
        np.testing.assert_array_equal(y.get(), x.get())
    assert b.state.step == a.state.step and b.state.t == a.state.t


@pytest.mark.parametrize("method", ["rk4-fdm-hip", "split-step-fft-hip"])
def test_resumed_noisy_run_is_bitwise_the_uninterrupted_run(tmp_path, method):
    """Wigner noise (the counter continues from the restored step), steady-state checks and
    g2 / momentum accumulators (their FP64 sums continue from the restored ones)."""
    def cfg_fn():
        cfg = quickstart_config(solver__method=method, grid__nx=64, grid__ny=48, grid__lx=40.0, grid__ly=30.0,
                                solver__dt=5e-3, solver__total_time=0.2, laser__P0=0.2, laser__sigma_space=8.0,
                                noise__stochastic="wigner", stop__steady_state=True, stop__check_every=0.05)
        cfg.noise.seed = 11
        cfg.result.accumulators = {"g2": {"every": 0.025}, "momentum-distribution": {"every": 0.05}}
        return cfg
    a, b = _pair(tmp_path, cfg_fn)
    _equal(a, b)
    np.testing.assert_array_equal(b.stop_times, a.stop_times)
    for name in ("density_mean", "density_squared_mean", "momentum_distribution"):
        np.testing.assert_array_equal(b.accumulated[name].value, a.accumulated[name].value)


def test_resumed_adaptive_run_continues_the_trial_sequence(tmp_path):
    def cfg_fn():
        return quickstart_config(solver__method="ifrk-fft-hip", solver__tableau="dp5", solver__adaptive=True,
                                 solver__rtol=1e-7, solver__atol=1e-9, grid__nx=32, grid__ny=32, grid__lx=30.0,
                                 grid__ly=30.0, solver__dt=0.5, solver__total_time=4.0, laser__P0=0.3,
                                 laser__sigma_space=6.0)
    a, b = _pair(tmp_path, cfg_fn)
    _equal(a, b)
    sa, sb = a.solver.step_controller.statistics, b.solver.step_controller.statistics
    assert sa == sb and sa["accepted"] >= 4


def test_resume_rejects_a_different_configuration(tmp_path):
    from polarism_amd import SimulationController
    cfg = quickstart_config(grid__nx=32, grid__ny=32, solver__total_time=0.01)
    cfg.checkpoint.path = str(tmp_path / "c.npz")
    SimulationController(cfg).run()
    other = quickstart_config(grid__nx=32, grid__ny=32, solver__total_time=0.02, solver__dt=2e-3)
    other.checkpoint.resume_from = str(tmp_path / "c.npz")
    with pytest.raises(ValueError, match="dt mismatch"):
        SimulationController(other)
    other = quickstart_config(grid__nx=32, grid__ny=32, solver__total_time=0.02, reservoir__reservoir_type="double")
    other.checkpoint.resume_from = str(tmp_path / "c.npz")
    with pytest.raises(ValueError, match="reservoir type mismatch"):
        SimulationController(other)
