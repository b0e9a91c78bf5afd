"""Accumulator host logic: configuration checks, schedules, g1 / g2 (no GPU)."""
import numpy as np
import pytest

from polarism_amd.accumulators import available_accumulators, g1, g2
from polarism_amd.config import ConfigValidationError
from polarism_amd.controller import _next_sample
from polarism_amd.validation import validate_config
from tests.helpers import quickstart_config


def test_registered_accumulators():
    assert {"momentum-distribution", "g2"} <= set(available_accumulators)


@pytest.mark.parametrize("acc,match", [({"nope": {}}, "Unknown accumulator 'nope'"),
                                       ({"g2": 3}, "parameter mapping"),
                                       ({"g2": {"every": 0.1, "segment": 8}}, "unknown keys"),
                                       ([("g2", {})], "must map names")])
def test_accumulator_config_is_checked(acc, match):
    cfg = quickstart_config(result__accumulators=acc)
    with pytest.raises(ConfigValidationError, match=match):
        validate_config(cfg)


def test_valid_accumulator_config_passes():
    validate_config(quickstart_config(result__accumulators={"g2": {"start": 0.5, "every": 0.1},
                                                            "momentum-distribution": {}}))


def test_next_sample_matches_brute_force():
    for start in (0, 3, 7):
        for every in (1, 2, 5):
            for stop in (0, 4, 20):
                due_steps = [s for s in range(start, stop + 1) if (s - start) % every == 0]
                for step in range(0, 25):
                    later = [s for s in due_steps if s > step]
                    assert _next_sample(step, start, every, stop) == (later[0] if later else None)


def test_g1_of_one_mode_is_a_unit_phase_and_g2_of_a_coherent_field_is_one():
    """A single occupied mode k0 has |g1(r)| = 1 everywhere with phase k0 . r (after the
    centring shift); a uniform coherent density has g2 = 1 and zero density gives 0."""
    n = 16
    nk = np.zeros((n, n))
    nk[n // 2 + 3, n // 2 - 2] = 5.0          # shifted spectrum: kx index -2, ky index +3
    c = g1(nk)
    np.testing.assert_allclose(np.abs(c), 1.0, atol=1e-12)
    y, x = np.meshgrid(np.arange(n) - n // 2, np.arange(n) - n // 2, indexing="ij")
    np.testing.assert_allclose(c, np.exp(2j * np.pi * (3 * y - 2 * x) / n), atol=1e-12)
    mean = np.full((4, 4), 2.0)
    np.testing.assert_allclose(g2(mean, mean ** 2), 1.0)
    assert np.all(g2(np.zeros((2, 2)), np.ones((2, 2))) == 0)
