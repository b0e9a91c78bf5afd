"""Profile pumps through the HIP path: rk4-fdm-hip with lattice and ring profile lasers against
the oracle run with the oracle's own restatement of the profile (builtin_profiles.py:66-74,
:200-240), complex128 at 1e-10."""
import numpy as np
import pytest

from oracle import gpe_oracle as O
from tests.helpers import oracle_physics, quickstart_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("profile", [
    {"type": "lattice", "lattice": "honeycomb", "constant": 9.0, "width": 2.0, "disorder": 0.15, "seed": 4,
     "defects": [[0, 0, 1]], "rotation": 10.0},
    {"type": "lattice", "lattice": "kagome", "constant": 7.0, "width": 1.5, "well": "supergaussian", "order": 2.0},
    {"type": "ring", "radius": 15.0, "width": 3.0},
])
def test_profile_pump_rk4_matches_oracle(profile):
    from polarism_amd import SimulationController
    cfg = quickstart_config(grid__nx=64, grid__ny=64, grid__lx=60.0, grid__ly=60.0, laser__laser_type="profile",
                            laser__P0=0.06, solver__dt=0.01, solver__total_time=1.0, reservoir__reservoir_type="single")
    cfg.laser.profile = dict(profile)
    ctl = SimulationController(cfg)
    ctl.run()
    g = O.Grid(64, 64, 60.0, 60.0)
    if profile["type"] == "ring":
        env = O.ring_profile(g, profile["radius"], profile["width"])
    else:
        kw = {k: v for k, v in profile.items() if k not in ("type", "lattice", "constant", "width")}
        kw["defects"] = [tuple(d) for d in kw.get("defects", [])]
        env = O.lattice_profile(g, profile["lattice"], profile["constant"], profile["width"], **kw)
    p = oracle_physics(cfg)
    psi = O.initial_psi(g.shape, cfg.noise.initial_amplitude, cfg.noise.seed).reshape(1, 1, 64, 64)
    res = (np.zeros((1, 1, 64, 64)),)
    for s in range(100):
        psi, res = O.rk4_fdm_step("single", p, g, psi, res, lambda t: env * 0.06, s * 0.01, 0.01)
    assert O.rel_l2(ctl.state.field.get(), psi) <= 1e-10
    assert O.rel_l2(ctl.reservoir.get_state()[0].get(), res[0]) <= 1e-10
