"""Result frames of a controller run written through the device snapshot and the writer
thread: every saved quantity equals what the reference's result nodes compute from the
saved field (simulation_controller.py:63-113, 477-576), times follow result.save_every, and
recording leaves the trajectory unchanged bit for bit."""
import numpy as np
import pytest

from tests.helpers import quickstart_config

pytestmark = pytest.mark.gpu


def _density(psi):
    # |psi|^2 as re*re + im*im in the field's precision (the frame kernel's form)
    return psi.real * psi.real + psi.imag * psi.imag


def _cfg(tmp_path, **kw):
    cfg = quickstart_config(grid__nx=64, grid__ny=48, grid__lx=40.0, grid__ly=30.0, solver__dt=5e-3,
                            solver__total_time=0.1, solver__precision="single", laser__P0=0.2,
                            laser__sigma_space=8.0, noise__stochastic="wigner", **kw)
    cfg.noise.seed = 11
    cfg.result.save_results = True
    cfg.result.save_every = 0.025
    cfg.result.save_psi = True
    cfg.result.output_directory = str(tmp_path)
    return cfg


def test_scalar_frames_npy_and_json(tmp_path):
    from polarism_amd import SimulationController
    from polarism_amd.results import load_results
    cfg = _cfg(tmp_path)
    cfg.result.formats = ("npy", "json")
    ctl = SimulationController(cfg)
    ctl.run()
    r = load_results(ctl.run_directory)
    np.testing.assert_allclose(r.time, [0.0, 0.025, 0.05, 0.075, 0.1], atol=1e-12)
    psi = r.fields["psi"]
    assert psi.shape == (5, 48, 64) and psi.dtype == np.complex64
    np.testing.assert_array_equal(psi[-1], ctl.state.psi)
    np.testing.assert_array_equal(r.fields["density"], _density(psi))
    np.testing.assert_allclose(r.fields["density"], np.abs(psi) ** 2, rtol=2e-6, atol=0)
    cell = ctl.grid.dx * ctl.grid.dy
    np.testing.assert_allclose(r.scalars["norm"], r.fields["density"].astype(np.float64).sum(axis=(1, 2)) * cell,
                               rtol=1e-12)
    n_r = ctl.reservoir.get_state()[0].get()[0, 0]
    np.testing.assert_array_equal(r.fields["N_R"][-1], n_r)
    np.testing.assert_array_equal(r.scalars["N_R_max"], r.fields["N_R"].max(axis=(1, 2)))
    np.testing.assert_allclose(r.static["pump"], ctl.pump(0.0), rtol=1e-7)
    assert r.scalar_groups["P_lasers"]["L0"].shape == (5,)
    assert r.attrs["datasets"]["fields/N_R"]["dims"] == ["time", "y", "x"]
    # the JSON backend holds the same frames
    j = load_results(_json_copy(ctl.run_directory, tmp_path))
    np.testing.assert_array_equal(j.fields["density"], r.fields["density"])
    np.testing.assert_array_equal(j.fields["psi"], psi)
    np.testing.assert_array_equal(j.scalars["norm"], r.scalars["norm"])
    # recording does not perturb the run
    plain = _cfg(tmp_path)
    plain.result.save_results = False
    other = SimulationController(plain)
    other.run()
    np.testing.assert_array_equal(other.state.field.get(), ctl.state.field.get())


def _json_copy(run_dir, tmp_path):
    """The run directory holds both layouts; the reader prefers NPY, so read JSON from a copy
    without the NPY batches."""
    import shutil
    dst = tmp_path / "json_only"
    dst.mkdir()
    for p in run_dir.iterdir():
        if p.suffix == ".json":
            shutil.copy(p, dst / p.name)
    return dst


def test_batched_spinor_frames(tmp_path):
    from polarism_amd import SimulationController
    from polarism_amd.results import load_results
    cfg = _cfg(tmp_path, physics__spin="spinor", physics__te_tm=0.05, physics__zeeman=0.02)
    cfg.result.formats = ("npy",)
    cfg.reservoir.reservoir_type = "double"
    cfg.reservoir.components = 2
    ctl = SimulationController(cfg, members=[{"laser.P0": 0.2}, {"laser.P0": 0.3}, {"laser.P0": 0.25}])
    ctl.run()
    r = load_results(ctl.run_directory)
    psi = r.fields["psi"]
    assert psi.shape == (5, 3, 2, 48, 64)
    np.testing.assert_array_equal(psi[-1], ctl.state.field.get())
    dens = _density(psi)
    np.testing.assert_array_equal(r.fields["density"], dens)
    np.testing.assert_array_equal(r.fields["density_plus"], dens[:, :, 0])
    np.testing.assert_array_equal(r.fields["density_minus"], dens[:, :, 1])
    p, m = psi[:, :, 0], psi[:, :, 1]
    coh = np.conj(p) * m
    stokes = np.stack([np.abs(p) ** 2 + np.abs(m) ** 2, 2 * coh.real, 2 * coh.imag,
                       np.abs(p) ** 2 - np.abs(m) ** 2], axis=2)
    scale = np.abs(stokes).max()
    np.testing.assert_allclose(r.fields["stokes"], stokes, rtol=0, atol=2e-6 * scale)
    s0 = stokes[:, :, 0]
    pol = np.sqrt((stokes[:, :, 1:] ** 2).sum(axis=2))
    dop = np.where(s0 > 0, pol / np.maximum(s0, 1e-30), 0)
    np.testing.assert_allclose(r.fields["degree_of_polarization"], dop, rtol=0, atol=1e-4)
    cell = ctl.grid.dx * ctl.grid.dy
    np.testing.assert_allclose(r.scalars["norm"], dens.astype(np.float64).sum(axis=(2, 3, 4)) * cell, rtol=1e-12)
    assert r.scalars["norm"].shape == (5, 3)
    n_a = ctl.reservoir.get_state()[0].get()
    np.testing.assert_array_equal(r.fields["nA"][-1], n_a)
    np.testing.assert_array_equal(r.scalars["nA_max"], r.fields["nA"].max(axis=(2, 3, 4)))
    assert r.attrs["members"] == 3
    assert r.attrs["datasets"]["fields/nA"]["dims"] == ["time", "member", "component", "y", "x"]
    # per-member laser maxima follow the members' P0 (same envelope shape)
    L0 = r.scalar_groups["P_lasers"]["L0"]
    assert L0.shape == (5, 3)
    np.testing.assert_allclose(L0[0] / L0[0][0], [1.0, 1.5, 1.25], rtol=1e-6)


def test_steady_state_stop_records_the_last_frame_and_stop_times(tmp_path):
    from polarism_amd import SimulationController
    from polarism_amd.results import load_results
    cfg = _cfg(tmp_path, stop__steady_state=True, stop__check_every=0.01, stop__rtol=1.0)
    cfg.noise.stochastic = "none"
    cfg.result.formats = ("npy",)
    cfg.result.save_every = 0.05
    cfg.result.accumulators = {"g2": {"every": 0.01}}
    ctl = SimulationController(cfg)
    ctl.run()
    r = load_results(ctl.run_directory)
    assert np.isfinite(ctl.stop_times).all()
    assert r.attrs["stop_time"] == ctl.stop_times.tolist()
    np.testing.assert_allclose(r.time[-1], ctl.state.t, atol=1e-12)
    np.testing.assert_array_equal(r.fields["psi"][-1], ctl.state.psi)
    np.testing.assert_array_equal(r.accumulated["g2"], ctl.accumulated["g2"].value)
    assert r.accumulated_meta["g2"]["samples"] > 0
