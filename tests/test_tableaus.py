"""Butcher tableaus: rooted-tree order conditions and agreement with an independent source.

The product's table (polarism_amd/tableaus.py) and the oracle's (SciPy's
published Dormand-Prince / DOP853 coefficients) are each checked against the
order conditions (reference tests/unit/test_tableaus.py:76-102), and against
each other.
"""
import itertools

import numpy as np
import pytest

from oracle import gpe_oracle as O
from polarism_amd.tableaus import TABLEAUS

# (elementary weight function, 1/gamma) for all rooted trees up to order 5
def _conditions(A, c, order):
    out = []
    e = np.ones_like(c)
    Ac = A @ c
    Ac2 = A @ c ** 2
    AAc = A @ Ac
    if order >= 1:
        out.append((e, 1.0))
    if order >= 2:
        out.append((c, 1 / 2))
    if order >= 3:
        out += [(c ** 2, 1 / 3), (Ac, 1 / 6)]
    if order >= 4:
        out += [(c ** 3, 1 / 4), (c * Ac, 1 / 8), (Ac2, 1 / 12), (AAc, 1 / 24)]
    if order >= 5:
        out += [(c ** 4, 1 / 5), (c ** 2 * Ac, 1 / 10), (c * Ac2, 1 / 15), (c * AAc, 1 / 30), (Ac * Ac, 1 / 20),
                (A @ c ** 3, 1 / 20), (A @ (c * Ac), 1 / 40), (A @ Ac2, 1 / 60), (A @ AAc, 1 / 120)]
    return out


@pytest.mark.parametrize("name,p,q", [("rk4", 4, None), ("dp5", 5, 4), ("dop853", 5, 5)])
def test_product_tableau_order_conditions(name, p, q):
    A, b, c, bh, bh3 = TABLEAUS[name].arrays()
    assert np.allclose(A.sum(axis=1), c, atol=1e-14)
    assert np.all(np.triu(A) == 0)
    for phi, target in _conditions(A, c, p):
        assert abs(b @ phi - target) < 1e-13
    if q:
        for phi, target in _conditions(A, c, q):
            assert abs(bh @ phi - target) < 1e-13
    if name == "dop853":
        for phi, target in _conditions(A, c, 3):
            assert abs(bh3 @ phi - target) < 1e-13


@pytest.mark.parametrize("name", ["rk4", "dp5", "dop853"])
def test_product_matches_independent_tables(name):
    A, b, c, bh, bh3 = TABLEAUS[name].arrays()
    t = O.oracle_tableau(name)
    assert np.allclose(A, t.A, atol=1e-15) and np.allclose(b, t.b, atol=1e-15) and np.allclose(c, t.c, atol=1e-15)
    if bh is not None:
        # embedded weights may differ by the sign convention of b - b_hat; the error norm uses |.|
        assert np.allclose(np.abs(b - bh), np.abs(t.b - t.b_hat), atol=1e-15)
    if bh3 is not None:
        assert np.allclose(np.abs(b - bh3), np.abs(t.b - t.b_hat3), atol=1e-15)


def test_embedded_orders_and_stage_counts():
    assert (TABLEAUS["rk4"].stages, TABLEAUS["dp5"].stages, TABLEAUS["dop853"].stages) == (4, 7, 12)
    assert TABLEAUS["dp5"].embedded_order == 4 and TABLEAUS["dop853"].embedded_order == 5
    assert TABLEAUS["rk4"].embedded_order is None
