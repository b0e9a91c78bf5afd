"""What the trained surrogate gains: warm-started members reach the same steady state sooner."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "polarism_amd", "surrogate", "weights", "fno_steady_256.npz")


def test_warm_start_reaches_the_cold_steady_state_sooner():
    """16 held-out members (a seed the training set never used): the warm start must reach the
    reference steady-state criterion (rtol 1e-6 per 1 ps) no later in total simulated time than
    the noise-seeded cold start, and land on the same density (rel-L2 <= 1e-3) where both converge."""
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
    print(f"converged cold {np.isfinite(cold.stop_times).sum()} warm {np.isfinite(hot.stop_times).sum()}; "
          f"median stop cold {np.nanmedian(cold.stop_times):.0f} ps warm {np.nanmedian(hot.stop_times):.0f} ps; "
          f"steps cold {cold.steps_run} warm {hot.steps_run}")
    assert np.isfinite(hot.stop_times).sum() >= np.isfinite(cold.stop_times).sum()
    assert np.nansum(np.where(both, hot.stop_times, 0)) < np.nansum(np.where(both, cold.stop_times, 0))
    agree = []
    for i in np.flatnonzero(both):
        dc = np.abs(cold.final_psi[i]) ** 2
        dw = np.abs(hot.final_psi[i]) ** 2
        nc, nw = np.linalg.norm(dc), np.linalg.norm(dw)
        # both at the zero steady state (norm below the criterion's 1e-30): they agree
        agree.append(0.0 if nc < 1e-30 and nw < 1e-30 else np.linalg.norm(dw - dc) / max(nc, 1e-30))
    agree = np.asarray(agree)
    print("density rel-L2 warm vs cold:", np.round(agree, 6))
    assert both.sum() >= max(4, int(0.8 * np.isfinite(cold.stop_times).sum()))
    assert np.median(agree) <= 1e-3


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
