"""The oracle's Lawson IFRK against the reference's golden trajectories.

tests/golden/ifrk4_reference.npz is the data file the reference keeps at
tests/data/ifrk4_reference/reference.npz (recipe: generate.py:17-59 there):
16x16 grid, 5 steps of dt 0.007, complex potential, pulsed Gaussian pump,
precision x reservoir (single, quadratic-double) x eta (0, 0.04).  Tolerances
are the reference's own (test_ifrk_reference.py:19-21).
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import gpe_oracle as O

GOLDEN = Path(__file__).parent / "golden" / "ifrk4_reference.npz"


def fixture_case(precision, kind, eta):
    g = O.Grid(16, 16, 20.0, 20.0)
    p = O.Physics(kinetic_relaxation_eta=eta)
    real = np.float32 if precision == "single" else np.float64
    cplx = np.complex64 if precision == "single" else np.complex128
    x, y = g.X, g.Y
    psi = np.asarray(0.03 * (1 + 0.1 * np.cos(2 * np.pi * x / 20)) * np.exp(1j * 0.2 * np.sin(2 * np.pi * y / 20)),
                     dtype=cplx).reshape(1, 1, 16, 16)
    nf = 1 if kind == "single" else 2
    res = tuple(np.asarray(0.2 + 0.02 * np.cos(2 * np.pi * x / 20) + i * 0.04, dtype=real).reshape(1, 1, 16, 16)
                for i in range(nf))
    pot = np.asarray(0.01 * np.cos(2 * np.pi * x / 20) - 0.003j * (1 + np.sin(2 * np.pi * y / 20)), dtype=cplx)
    env = O.gaussian_profile(g, 15.0).astype(real)
    pump_at = lambda t: env * O.pulse_gaussian_coeff(t, 0.4, 1.2, 0.03, 3.0, 0.6)
    return g, p, psi, res, pot, pump_at


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("kind", ["single", "quadratic-double"])
@pytest.mark.parametrize("eta", [0.0, 0.04])
def test_oracle_lawson_rk4_matches_reference_trajectories(precision, kind, eta):
    g, p, psi, res, pot, pump_at = fixture_case(precision, kind, eta)
    prob = O.LawsonProblem(kind, p, g, precision=precision, potential=pot, tableau="rk4")
    for s in range(5):
        _, _, psi, res = prob.step(psi, res, pump_at, s * 0.007, 0.007)
    key = f"{precision}_{kind}_eta{eta:g}"
    tol = 1e-5 if precision == "single" else 1e-12
    with np.load(GOLDEN) as ref:
        np.testing.assert_allclose(psi[0, 0], ref[key + "_psi"], rtol=tol,
                                   atol=tol * max(1e-3, np.max(abs(ref[key + "_psi"]))))
        for i, r in enumerate(res):
            np.testing.assert_allclose(r[0, 0], ref[f"{key}_res{i}"], rtol=tol,
                                       atol=tol * max(1e-3, np.max(abs(ref[f"{key}_res{i}"]))))


def test_golden_fixture_is_complete():
    with np.load(GOLDEN) as ref:
        assert len(ref.files) == 20
