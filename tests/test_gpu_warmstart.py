"""What the trained surrogate gains: warm-started members reach the cold start's steady state sooner.

The cold (noise-seeded) run of each member is the judge: a warm member counts only where both runs
meet the steady-state criterion and their final densities agree within 1e-3 relative L2."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "polarism_amd", "surrogate", "weights", "fno_steady_256.npz")


def _density_error(a, b):
    da, db = np.abs(a) ** 2, np.abs(b) ** 2
    den = np.linalg.norm(db)
    if den < 1e-30 and np.linalg.norm(da) < 1e-30:
        return 0.0
    return float(np.linalg.norm(da - db) / max(den, 1e-30))


def test_warm_start_reaches_the_cold_steady_state_sooner():
    """16 held-out members (a seed the training set never used), rtol 1e-6 per 1 ps:
    * every member whose warm and cold runs both converge lands within 1e-3 of the cold density
      (the gate, on the maximum);
    * every member predicted to hold no condensate (psi = 0 start) is one whose cold run decayed
      to the zero state, where that cold run converged;
    * on the members passing the gate, the warm runs need less simulated time in total and the
      last of them converges sooner in wall time than the last cold one."""
    from polarism_amd.surrogate import warmstart as ws
    from polarism_amd.sweep import PackedSweep
    if not os.path.exists(WEIGHTS):
        pytest.fail("trained weights are missing")
    members = ws.make_members(16, seed=424242)
    params, cfg, norm = ws.load_weights(WEIGHTS)
    y = ws.predict(params, ws.features(members), cfg)
    warm = ws.warm_start(members, y, norm)
    runs = {}
    for name, ms in (("cold", members), ("warm", warm)):
        sw = PackedSweep(ms, reservoir="single", precision="single", dt=0.05, check_every=1.0, rtol=1e-6,
                         max_time=3000.0)
        runs[name] = sw.run()
    cold, hot = runs["cold"], runs["warm"]
    both = np.isfinite(cold.stop_times) & np.isfinite(hot.stop_times)
    err = np.full(16, np.nan)
    for i in np.flatnonzero(both):
        err[i] = _density_error(hot.final_psi[i], cold.final_psi[i])
    zero = np.asarray([m.label["warm"] == "zero" for m in warm])
    print(f"converged cold {np.isfinite(cold.stop_times).sum()} warm {np.isfinite(hot.stop_times).sum()}; "
          f"stop cold {cold.stop_times} warm {hot.stop_times}; zero starts {np.flatnonzero(zero)}")
    print("density rel-L2 warm vs cold:", np.round(err, 6))
    assert both.sum() >= 6
    assert np.nanmax(err) <= 1e-3
    cold_zero = np.asarray([np.linalg.norm(np.abs(p) ** 2) < 1e-30 for p in cold.final_psi])
    assert np.all(cold_zero[zero & both])
    gated = both & (err <= 1e-3)
    assert hot.stop_times[gated].sum() < cold.stop_times[gated].sum()
    assert np.max(hot.wall_at_stop[gated]) < np.max(cold.wall_at_stop[gated])


def test_device_pipeline_equals_host_pipeline():
    """predict_states (features, forward and denormalisation on the GPU) gives the warm starts
    of predict + warm_start on the host: the same zero-state choices and fields to rounding."""
    from polarism_amd.surrogate import warmstart as ws
    members = ws.make_members(16, seed=99)
    params, cfg, norm = ws.load_weights(WEIGHTS)
    host = ws.warm_start(members, ws.predict(params, ws.features(members), cfg), norm)
    dev = ws.warm_start(members, None, norm, states=ws.predict_states(params, members, cfg, norm))
    for a, b in zip(host, dev):
        assert a.label["warm"] == b.label["warm"]
        np.testing.assert_allclose(b.psi0, a.psi0, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(b.res0[0], a.res0[0], rtol=1e-5, atol=1e-6)
