"""Quadrature tables used by the JVIE T-matrix path (oracle side).

TEST INFRASTRUCTURE ONLY: nothing in the product imports this module.

The tables are published rules, typed here as data in the form the
reference uses them (reference-relative coordinates, weights summing to the
reference element's measure):

* tetrahedron, 1 point   -- src/integration_points.f90:315-323 (inttetra order 1)
* tetrahedron, 24 points -- src/integration_points.f90:355-466 (inttetra order 5,
  a degree-6 symmetric rule: 4+4+4+12 orbits)
* triangle, 1 point      -- src/integration_points.f90:138-144 (inttri order 1;
  the centroid is typed as 3.333333333333330d-01, not 1/3, and kept so)
* triangle, 7 points     -- src/integration_points.f90:165-187 (inttri order 5,
  the degree-5 Radon/Strang-Fix 7-point rule)
* cube sampling points   -- src/integration_points.f90:503-805: the 98 points are
  exactly the boundary points of the lattice {-1,-1/2,0,1/2,1}^3.

Pinning: tests/test_oracle_quadrature.py integrates monomials exactly
(degree 6 on the tet rule, degree 5 on the triangle rule) and checks the
98-point set, so a mistyped digit fails a test.
"""
import numpy as np

TET1_P = np.array([[0.25, 0.25, 0.25]])
TET1_W = np.array([1.666666666666667e-01])

TRI1_P = np.array([[3.333333333333330e-01, 3.333333333333330e-01]])
TRI1_W = np.array([0.5])

_a = 3.561913862225449e-01
_b = 2.146028712591517e-01
_c = 8.779781243961660e-01
_d = 4.067395853461134e-02
_e = 3.298632957317306e-02
_f = 3.223378901422757e-01
_g = 2.696723314583159e-01
_h = 6.366100187501753e-02
_i = 6.030056647916491e-01
TET24_P = np.array([
    [_a, _b, _b], [_b, _b, _b], [_b, _b, _a], [_b, _a, _b],
    [_c, _d, _d], [_d, _d, _d], [_d, _d, _c], [_d, _c, _d],
    [_e, _f, _f], [_f, _f, _f], [_f, _f, _e], [_f, _e, _f],
    [_g, _h, _h], [_h, _g, _h], [_h, _h, _g], [_i, _h, _h],
    [_h, _i, _h], [_h, _h, _i], [_h, _g, _i], [_g, _i, _h],
    [_i, _h, _g], [_h, _i, _g], [_g, _h, _i], [_i, _g, _h],
])
TET24_W = np.array([6.653791709694644e-03] * 4 + [1.679535175886776e-03] * 4
                   + [9.226196923942397e-03] * 4 + [8.035714285714281e-03] * 12)

TRI7_P = np.array([
    [3.333333333333330e-01, 3.333333333333330e-01],
    [5.971587178977000e-02, 4.701420641051150e-01],
    [4.701420641051150e-01, 4.701420641051150e-01],
    [4.701420641051150e-01, 5.971587178977000e-02],
    [7.974269853530870e-01, 1.012865073234560e-01],
    [1.012865073234560e-01, 1.012865073234560e-01],
    [1.012865073234560e-01, 7.974269853530870e-01],
])
TRI7_W = np.array([1.125000000000000e-01] + [6.619707639425300e-02] * 3
                  + [6.296959027241350e-02] * 3)


def cube98_points():
    """Boundary points of the 5x5x5 lattice on [-1,1]^3 (98 points).

    src/integration_points.f90:503-805 lists the same set (order differs;
    the least-squares fit that uses them does not depend on the order).
    """
    g = np.array([-1.0, -0.5, 0.0, 0.5, 1.0])
    pts = []
    for z in g:
        for y in g:
            for x in g:
                if max(abs(x), abs(y), abs(z)) == 1.0:
                    pts.append((x, y, z))
    return np.array(pts)


CUBE98_P = cube98_points()
