"""The Sweep entry point's host logic: override paths, member order, batch planning and the
sweep.json manifest (sweep.py:98-278, sweep_config.py:18-124).  No GPU."""
import json

import pytest

from polarism_amd.sweep import Sweep, override_kind
from tests.helpers import quickstart_config


def test_override_kinds():
    assert override_kind("physics.g_C") == "batch"
    assert override_kind("laser.P0") == "batch"
    assert override_kind("laser.profile.radius") == "batch"
    assert override_kind("laser.profile.type") == "group"
    assert override_kind("physics.spin") == "group"
    assert override_kind("noise.seed") == "batch"
    assert override_kind("noise.stochastic") == "group"
    for path in ("grid.nx", "solver.dt", "reservoir.reservoir_type", "stop.rtol"):
        assert override_kind(path) == "group"
    for bad in ("physics.g_X", "nosection.x", "grid", "physics.g_C.sub"):
        with pytest.raises(ValueError):
            override_kind(bad)


def test_product_and_zip_member_order():
    cfg = quickstart_config()
    sw = Sweep(cfg, {"laser.P0": [0.01, 0.02], "physics.g_C": [1e-3, 2e-3, 3e-3]})
    assert [m["laser.P0"] for m in sw.members] == [0.01] * 3 + [0.02] * 3
    assert [m["physics.g_C"] for m in sw.members] == [1e-3, 2e-3, 3e-3] * 2
    sw = Sweep(cfg, {"laser.P0": [0.01, 0.02], "physics.g_C": [1e-3, 2e-3]}, mode="zip")
    assert sw.members == [{"laser.P0": 0.01, "physics.g_C": 1e-3}, {"laser.P0": 0.02, "physics.g_C": 2e-3}]
    with pytest.raises(ValueError):
        Sweep(cfg, {"laser.P0": [0.01], "physics.g_C": [1e-3, 2e-3]}, mode="zip")
    with pytest.raises(ValueError):
        Sweep(cfg, {"laser.P0": []})


def test_grid_sizes_pack_into_one_launch_for_rk4_fdm_hip():
    """Members of different grids share one packed batch; the reference would run three."""
    cfg = quickstart_config(laser__laser_type="continuous-gaussian")
    sw = Sweep(cfg, {"grid.nx": [64, 96, 128], "laser.P0": [0.02, 0.03]})
    assert len(sw.batches) == 1 and sw.batches[0]["packed"] and sw.batches[0]["size"] == 6


def test_non_packable_members_group_like_the_reference():
    cfg = quickstart_config(solver__method="split-step-fft-hip", laser__laser_type="continuous-gaussian")
    sw = Sweep(cfg, {"grid.nx": [64, 128], "laser.P0": [0.02, 0.03]})
    assert [b["packed"] for b in sw.batches] == [False, False]
    assert [b["indices"] for b in sw.batches] == [[0, 1], [2, 3]]
    # a time-dependent pump keeps rk4-fdm-hip members on the controller path
    cfg = quickstart_config(laser__laser_type="pulse-gaussian")
    sw = Sweep(cfg, {"laser.P0": [0.2, 0.3]})
    assert [b["packed"] for b in sw.batches] == [False]


def test_manifest_maps_every_member():
    cfg = quickstart_config(laser__laser_type="continuous-gaussian")
    sw = Sweep(cfg, {"grid.nx": [64, 96], "physics.g_C": [1e-3, 2e-3]})
    m = json.loads(json.dumps(sw.manifest(7, [7] * 4)))
    assert m["format_version"] == 1 and m["mode"] == "product" and m["seed"] == 7
    assert m["parameter_keys"] == ["grid.nx", "physics.g_C"] and len(m["parameters"]) == 4
    assert m["batch_sizes"] == [4] and m["members"][3] == {"group": 0, "batch": 0, "position": 3}
    assert set(m["config_signature"]) >= {"grid", "solver", "physics", "laser", "stop"}
