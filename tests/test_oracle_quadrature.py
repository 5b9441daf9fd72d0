"""Pins the quadrature tables by exact monomial integration.

A mistyped digit in a point or weight breaks exactness at the rule's degree.
"""
import math

import numpy as np

from oracle import quadrature as q


def _tet_exact(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


def _tri_exact(a, b):
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


def test_tet24_degree6():
    P, W = q.TET24_P, q.TET24_W
    assert P.shape == (24, 3)
    for a in range(7):
        for b in range(7 - a):
            for c in range(7 - a - b):
                val = np.sum(W * P[:, 0] ** a * P[:, 1] ** b * P[:, 2] ** c)
                assert abs(val - _tet_exact(a, b, c)) < 2e-15, (a, b, c)


def test_tet24_not_degree7():
    # sanity: the rule is not exact beyond its degree (guards against a vacuous test)
    P, W = q.TET24_P, q.TET24_W
    errs = [abs(np.sum(W * P[:, 0] ** a * P[:, 1] ** (8 - a)) - _tet_exact(a, 8 - a, 0)) for a in range(9)]
    assert max(errs) > 1e-9


def test_tri7_degree5():
    P, W = q.TRI7_P, q.TRI7_W
    for a in range(6):
        for b in range(6 - a):
            val = np.sum(W * P[:, 0] ** a * P[:, 1] ** b)
            # the centroid is typed as 0.333333333333330 in the reference table
            assert abs(val - _tri_exact(a, b)) < 1e-14, (a, b)


def test_one_point_rules():
    assert abs(q.TET1_W.sum() - 1.0 / 6.0) < 1e-16
    assert abs(q.TRI1_W.sum() - 0.5) == 0.0
    assert np.allclose(q.TET1_P, 0.25)


def test_cube98_points():
    P = q.CUBE98_P
    assert P.shape == (98, 3)
    assert len({tuple(p) for p in P}) == 98
    assert np.all(np.max(np.abs(P), axis=1) == 1.0)
    assert set(np.unique(P)) == {-1.0, -0.5, 0.0, 0.5, 1.0}
