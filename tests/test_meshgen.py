"""The cube-split mesh generator: counts, orientation, conformity, rescaling."""
import numpy as np

from scadyn_amd import meshgen


def _faces(e):
    return np.sort(np.concatenate([e[:, [1, 2, 3]], e[:, [0, 2, 3]], e[:, [0, 1, 3]], e[:, [0, 1, 2]]]), axis=1)


def test_sphere_24_cells():
    c, e = meshgen.sphere(24, seed=0)
    assert e.shape == (43248, 4)          # centre-in rule on a 24^3 lattice
    cc = c[e]
    d = np.einsum('ni,ni->n', cc[:, 1] - cc[:, 0], np.cross(cc[:, 2] - cc[:, 0], cc[:, 3] - cc[:, 0]))
    assert (d > 0).all()
    _, cnt = np.unique(_faces(e), axis=0, return_counts=True)
    assert cnt.max() == 2                 # conforming: a face is shared by at most two tets
    assert abs(meshgen.tet_volumes(c, e).sum() - 7208 * (2.0 / 24) ** 3) < 1e-9   # 43248/6 cells


def test_jitter_is_seeded_and_interior_only():
    c1, e1 = meshgen.sphere(8, seed=0)
    c2, e2 = meshgen.sphere(8, seed=0)
    c0, _ = meshgen.sphere(8, jitter=0.0)
    assert np.array_equal(c1, c2) and np.array_equal(e1, e2)
    moved = np.any(c1 != c0, axis=1)
    assert 0 < moved.sum() < len(c1)
    h = 2.0 / 8
    assert np.abs(c1 - c0).max() <= 0.1 * h + 1e-15


def test_rescale_volume_equivalent_radius():
    c, e = meshgen.sphere(6)
    c2 = meshgen.rescale_to_radius(c, e, 1e-7)
    v = meshgen.tet_volumes(c2, e).sum()
    assert abs((3 * v / (4 * np.pi)) ** (1 / 3) - 1e-7) < 1e-20


def test_gaussian_ellipsoid_shape():
    c, e = meshgen.gaussian_ellipsoid(n_cells=16, seed=3)
    cc = c[e]
    d = np.einsum('ni,ni->n', cc[:, 1] - cc[:, 0], np.cross(cc[:, 2] - cc[:, 0], cc[:, 3] - cc[:, 0]))
    assert (d > 0).all()
    ext = c.max(axis=0) - c.min(axis=0)
    assert ext[0] > ext[1] > ext[2]       # semi-axes 1 : 0.8 : 0.6
