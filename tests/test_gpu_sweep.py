"""Packed sweeps: uneven grids in one launch, retirement, steady-state semantics."""
import numpy as np
import pytest

from oracle import gpe_oracle as O
from tests.helpers import quickstart_config

pytestmark = pytest.mark.gpu


def _members(sizes, reservoir="double", seed=5):
    from polarism_amd.config import PhysicsConstants
    from polarism_amd.pumps import random_pump, threshold
    from polarism_amd.sweep import SweepMember
    p = PhysicsConstants()
    rng = np.random.default_rng(seed)
    out = []
    for i, (nx, ny) in enumerate(sizes):
        lx, ly = 0.5 * nx, 0.5 * ny
        env, _ = random_pump(rng, nx, ny, lx, ly, "gaussians")
        out.append(SweepMember(nx, ny, lx, ly, p, env, rng.uniform(0.5, 3.0) * threshold(p, reservoir),
                               seed=100 + i))
    return out


def _run(members, reservoir="double", precision="single", max_time=2.0, rtol=0.0, tile_h=8):
    from polarism_amd.sweep import PackedSweep
    sw = PackedSweep(members, reservoir=reservoir, precision=precision, dt=0.05, check_every=0.5, rtol=rtol,
                     max_time=max_time, tile_h=tile_h)
    return sw.run()


SIZES = [(64, 64), (96, 96), (128, 96), (192, 64), (64, 128), (256, 256)]


def test_uneven_grid_members_equal_single_member_runs_bitwise():
    members = _members(SIZES)
    packed = _run(members)
    for i, m in enumerate(members):
        single = _run([m])
        np.testing.assert_array_equal(packed.final_psi[i], single.final_psi[0])
        for a, b in zip(packed.final_res[i], single.final_res[0]):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision,tol", [("double", 1e-10), ("single", 1e-5)])
def test_uneven_grid_member_matches_oracle(precision, tol):
    members = _members(SIZES[:3])
    res = _run(members, precision=precision, max_time=1.0)
    from polarism_amd.noise import make_initial_psi
    for i, m in enumerate(members):
        g = O.Grid(m.nx, m.ny, m.lx, m.ly)
        real = np.float32 if precision == "single" else np.float64
        psi = make_initial_psi((m.ny, m.nx), 1e-3, seed=m.seed, precision=precision).reshape(1, 1, m.ny, m.nx)
        res0 = (np.zeros((1, 1, m.ny, m.nx), real), np.zeros((1, 1, m.ny, m.nx), real))
        env = m.envelope.astype(real)
        p = O.Physics()
        for s in range(20):
            psi, res0 = O.rk4_fdm_step("double", p, g, psi, res0, lambda t: env * m.P0, s * 0.05, 0.05)
        assert O.rel_l2(res.final_psi[i], psi[0, 0]) <= tol
        assert O.rel_l2(res.final_res[i][0], res0[0][0, 0]) <= tol


def test_retirement_does_not_change_running_members():
    members = _members(SIZES[:4], reservoir="single")
    # member 0 gets no pump and a zero initial field: it converges (norm < 1e-30) at the second check
    members[0].P0 = 0.0
    members[0].psi0 = np.zeros((64, 64), np.complex64)
    a = _run(members, reservoir="single", max_time=3.0, rtol=1e-12)
    assert np.isfinite(a.stop_times[0]) and not np.isfinite(a.stop_times[1:]).any()
    b = _run(members[1:], reservoir="single", max_time=3.0, rtol=1e-12)
    for i in range(1, 4):
        np.testing.assert_array_equal(a.final_psi[i], b.final_psi[i - 1])


def test_whole_step_sweep_equals_stage_sweep_bitwise():
    members = _members(SIZES[:5], reservoir="double")
    from polarism_amd.sweep import PackedSweep
    runs = []
    for form, tile_h in (("stage", 16), ("step", 32), ("strip", 24)):
        sw = PackedSweep(members, reservoir="double", dt=0.05, check_every=0.5, rtol=1e-3, max_time=3.0,
                         tile_h=tile_h, form=form)
        runs.append(sw.run())
    a = runs[0]
    for b in runs[1:]:
        np.testing.assert_array_equal(a.stop_times, b.stop_times)
        for i in range(len(members)):
            np.testing.assert_array_equal(a.final_psi[i], b.final_psi[i])
            for x, y in zip(a.final_res[i], b.final_res[i]):
                np.testing.assert_array_equal(x, y)


def test_autotuner_picks_a_valid_form_and_caches_it(tmp_path):
    from polarism_amd.autotune import Autotuner, candidates, tune_rk4
    from polarism_amd.layout import MemberSpec
    from polarism_amd.config import PhysicsConstants
    p = PhysicsConstants()
    specs = [MemberSpec(nx=n, ny=n, dx=0.5, dy=0.5, physics=p, reservoir="single", dt=0.05) for n in (64, 128, 96)]
    tuner = Autotuner(tmp_path / "tune.json")
    choice = tune_rk4(specs, "single", "single", steps=4, tuner=tuner)
    assert choice in candidates("single", 1)
    assert (tmp_path / "tune.json").exists()
    assert tune_rk4(specs, "single", "single", steps=4, tuner=Autotuner(tmp_path / "tune.json")) == choice


def test_reference_sweep_packs_grid_sizes_and_matches_single_runs(tmp_path):
    """Sweep(cfg, params).run() with grid.nx and laser.P0 varied: one packed launch for all four
    members (two grid shapes), each member equal to its own SimulationController run; the run
    directory holds sweep.json, and a resume reads the completed batch back unchanged."""
    import json
    from polarism_amd import SimulationController
    from polarism_amd.controller import apply_overrides
    from polarism_amd.sweep import Sweep
    cfg = quickstart_config(grid__nx=32, grid__ny=32, grid__lx=24.0, grid__ly=24.0,
                            laser__laser_type="continuous-gaussian", laser__sigma_space=5.0, solver__dt=0.01,
                            solver__total_time=0.5, noise__seed=11)
    cfg.result.save_results = True
    cfg.result.output_directory = str(tmp_path)
    sw = Sweep(cfg, {"grid.nx": [32, 48], "laser.P0": [0.02, 0.05]})
    assert len(sw.batches) == 1 and sw.batches[0]["packed"]
    out = sw.run()
    for i, overrides in enumerate(sw.members):
        single_cfg = apply_overrides(cfg, overrides)
        single_cfg.result.save_results = False
        single = SimulationController(single_cfg)
        single.run()
        assert O.rel_l2(out[i]["psi"], single.state.psi) <= 1e-12
        assert O.rel_l2(out[i]["reservoir"][0], single.reservoir.get_state()[0].get()[0, 0]) <= 1e-12
    manifest = json.loads((sw.run_directory / "sweep.json").read_text())
    assert manifest["batch_sizes"] == [4] and manifest["seed"] == 11
    again = Sweep(cfg, {"grid.nx": [32, 48], "laser.P0": [0.02, 0.05]}).run(resume=sw.run_directory)
    for a, b in zip(out, again):
        assert (a["psi"] == b["psi"]).all()
