"""Pins the singular integrals of the near zone against brute-force quadrature.

Brute force: split the triangle (tet) into sub-triangles (sub-tets) with apex at
the projection of the field point and integrate in Duffy coordinates, where the
1/R singularity is cancelled by the Jacobian, with a 96^2 (64^3) Gauss-Legendre
tensor rule.  Independent of the closed forms being tested.
"""
import numpy as np
import pytest

from oracle import pfft_ref as R

def _gl(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


GL_X, GL_W = _gl(96)      # triangles
GT_X, GT_W = _gl(64)      # tetrahedra


def duffy_tri(f, p0, p1, p2):
    """int over triangle (p0 apex, p1, p2) of f(r) dS."""
    u, v = np.meshgrid(GL_X, GL_X, indexing='ij')
    wu, wv = np.meshgrid(GL_W, GL_W, indexing='ij')
    pts = p0 + u[..., None] * ((p1 - p0) + v[..., None] * (p2 - p1))
    area2 = np.linalg.norm(np.cross(p1 - p0, p2 - p0))
    return np.sum(f(pts) * wu * wv * u * area2)


def tri_brute(f, tri, r):
    n = R.tri_normal(tri)
    rp = r - np.dot(r - tri[0], n) * n        # projection into the plane
    total = 0.0
    for a, b in ((0, 1), (1, 2), (2, 0)):
        # signed sub-triangle (rp, p_a, p_b): orientation relative to the face normal
        s = np.sign(np.dot(np.cross(tri[a] - rp, tri[b] - rp), n))
        if s != 0:
            total = total + s * duffy_tri(f, rp, tri[a], tri[b])
    return total


def duffy_tet(f, p0, p1, p2, p3):
    u, v, w = np.meshgrid(GT_X, GT_X, GT_X, indexing='ij')
    wu, wv, ww = np.meshgrid(GT_W, GT_W, GT_W, indexing='ij')
    pts = p0 + u[..., None] * ((p1 - p0) + v[..., None] * ((p2 - p1) + w[..., None] * (p3 - p2)))
    vol6 = abs(np.dot(p1 - p0, np.cross(p2 - p0, p3 - p0)))
    return np.sum(f(pts) * wu * wv * ww * u * u * v * vol6)


def tet_brute(f, tet, r):
    total = 0.0
    for face in R.FACES:
        tri = tet[face]
        s = np.sign(np.dot(np.cross(tri[1] - tri[0], tri[2] - tri[0]), tri[0] - r))
        if s != 0:
            total = total + s * duffy_tet(f, r, tri[0], tri[1], tri[2])
    return total


TRI = np.array([[0.1, 0.2, 0.0], [1.3, 0.1, 0.05], [0.4, 1.1, -0.1]]) * 1e-8
TET = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.1, 0.1], [0.1, 0.3, 0.9]]) * 1e-8


def field_points_tri():
    n = R.tri_normal(TRI)
    c = TRI.mean(axis=0)
    pts = [c + 0.3e-8 * n, c - 0.05e-8 * n,                 # above / below the centroid
           TRI[0] + 0.2 * (TRI[1] - TRI[0]) + 0.6e-9 * n,    # near an edge
           c + 2e-8 * n + 1e-8 * (TRI[1] - TRI[0]) / 1e-8,   # far
           c,                                                # in-plane, inside
           0.5 * (TRI[0] + TRI[1]) - 0.3 * (c - TRI[2]) * 0.5]  # in-plane, outside
    return pts


@pytest.mark.parametrize('ip', range(6))
def test_analytical_integration(ip):
    r = field_points_tri()[ip]
    n = R.tri_normal(TRI)
    K1, K2 = R.analytical_integration(TRI, n, r)

    def inv_r(p):
        return 1.0 / np.linalg.norm(p - r, axis=-1)

    def lin_r(p):
        return np.linalg.norm(p - r, axis=-1)
    b1 = tri_brute(inv_r, TRI, r)
    b2 = tri_brute(lin_r, TRI, r)
    assert abs(K1 - b1) / abs(b1) < 1e-11
    assert abs(K2 - b2) / abs(b2) < 1e-11


K = 2.0e7  # ka = 2 with a = 1e-7 m


@pytest.mark.parametrize('case', ['inside', 'vertex-near', 'outside'])
def test_int_V_G(case):
    c = TET.mean(axis=0)
    # field points stay more than 1e-10 m from every face plane: closer than that
    # the reference drops the |w0|*atan term (|w0| > 1e-10 test,
    # src/singularity_subtraction.f90:73-82), a deliberate part of the restatement
    r = {'inside': c, 'vertex-near': TET[0] + 0.05e-8, 'outside': c + np.array([2e-8, 0, 0])}[case]
    Pb, wb = R.linmap_tet(TET, R.quad.TET24_P, R.quad.TET24_W)
    got = R.int_V_G(r[None], K, Pb, wb, TET)[0]

    def green(p):
        d = np.linalg.norm(p - r, axis=-1)
        return np.exp(1j * K * d) / (4 * np.pi * d)
    ref = tet_brute(green, TET, r)
    # the smooth remainder is integrated with the 24-point rule
    assert abs(got - ref) / abs(ref) < 1e-8


@pytest.mark.parametrize('case', ['on-face', 'off-face'])
def test_int_S_G(case):
    n = R.tri_normal(TRI)
    c = TRI.mean(axis=0)
    r = c if case == 'on-face' else c + 0.4e-8 * n
    Pb, wb = R.linmap_tri(TRI, R.quad.TRI7_P, R.quad.TRI7_W)
    got = R.int_S_G(r[None], TRI, n, K, Pb, wb)[0]

    def green(p):
        d = np.linalg.norm(p - r, axis=-1)
        return np.exp(1j * K * d) / (4 * np.pi * d)
    ref = tri_brute(green, TRI, r)
    assert abs(got - ref) / abs(ref) < 1e-7


def test_coincident_point_limit():
    # R < cut uses the limit ik/(4 pi) of G - (1/R - k^2 R/2)/(4 pi)
    Rs = np.array([1e-14, 1e-13])
    lim = R._green_minus_static(Rs, K)
    assert np.allclose(lim, 1j * K / (4 * np.pi), rtol=1e-6, atol=0)
