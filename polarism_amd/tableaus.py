"""Explicit Runge-Kutta tableaus for the Lawson interaction-picture solver.

Same set of methods and roles as solver/tableaus.py:15-126 of the reference
(A, b, c, embedded b_hat and, for DOP853, the third-order b_hat3):

* ``rk4``    -- the classical fourth-order method (no embedded pair);
* ``dp5``    -- Dormand & Prince (1980) 5(4), FSAL, 7 stages, as exact rationals;
* ``dop853`` -- Hairer, Norsett & Wanner's DOP853 8(5,3), 12 stages; the
  coefficients are the published ones distributed with SciPy
  (``scipy.integrate._ivp.dop853_coefficients``), the source the reference
  also names.
"""
from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction as Q

import numpy as np


@dataclass(frozen=True)
class Tableau:
    name: str
    order: int
    embedded_order: int | None
    A: tuple
    b: tuple
    c: tuple
    b_hat: tuple | None = None
    b_hat3: tuple | None = None

    @property
    def stages(self) -> int:
        return len(self.b)

    def arrays(self):
        """float64 (A, b, c, b_hat or None, b_hat3 or None)."""
        f = lambda v: None if v is None else np.asarray([float(x) for x in v], dtype=np.float64)
        A = np.asarray([[float(x) for x in row] for row in self.A], dtype=np.float64)
        return A, f(self.b), f(self.c), f(self.b_hat), f(self.b_hat3)


def _rk4() -> Tableau:
    z = Q(0)
    A = ((z, z, z, z), (Q(1, 2), z, z, z), (z, Q(1, 2), z, z), (z, z, Q(1), z))
    return Tableau("rk4", 4, None, A, (Q(1, 6), Q(1, 3), Q(1, 3), Q(1, 6)), (z, Q(1, 2), Q(1, 2), Q(1)))


def _dp5() -> Tableau:
    z = Q(0)
    b = (Q(35, 384), z, Q(500, 1113), Q(125, 192), Q(-2187, 6784), Q(11, 84), z)
    A = (
        (z,) * 7,
        (Q(1, 5),) + (z,) * 6,
        (Q(3, 40), Q(9, 40)) + (z,) * 5,
        (Q(44, 45), Q(-56, 15), Q(32, 9)) + (z,) * 4,
        (Q(19372, 6561), Q(-25360, 2187), Q(64448, 6561), Q(-212, 729)) + (z,) * 3,
        (Q(9017, 3168), Q(-355, 33), Q(46732, 5247), Q(49, 176), Q(-5103, 18656), z, z),
        b[:6] + (z,),
    )
    c = (z, Q(1, 5), Q(3, 10), Q(4, 5), Q(8, 9), Q(1), Q(1))
    b_hat = (Q(5179, 57600), z, Q(7571, 16695), Q(393, 640), Q(-92097, 339200), Q(187, 2100), Q(1, 40))
    return Tableau("dp5", 5, 4, A, b, c, b_hat)


def _dop853() -> Tableau:
    from scipy.integrate._ivp import dop853_coefficients as d
    n = d.N_STAGES
    A = tuple(tuple(float(x) for x in d.A[i, :n]) for i in range(n))
    b = tuple(float(x) for x in d.B)
    c = tuple(float(x) for x in d.C[:n])
    b_hat = tuple(float(bi - ei) for bi, ei in zip(d.B, d.E5[:n]))
    b_hat3 = tuple(float(bi - ei) for bi, ei in zip(d.B, d.E3[:n]))
    return Tableau("dop853", 8, 5, A, b, c, b_hat, b_hat3)


class _Lazy(dict):
    _makers = {"rk4": _rk4, "dp5": _dp5, "dop853": _dop853}

    def __missing__(self, key):
        if key not in self._makers:
            raise KeyError(key)
        value = self._makers[key]()
        self[key] = value
        return value

    def __contains__(self, key):
        return key in self._makers

    def keys(self):
        return self._makers.keys()


TABLEAUS: dict[str, Tableau] = _Lazy()
