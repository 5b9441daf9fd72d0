"""Synthetic tetrahedral meshes for the JVIE T-matrix workloads.

scadyn reads its meshes from HDF5 files made by tetgen/pymesh
(other/scadyn_mesh.py in the reference); none of those tools exist here, so the
meshes are made by a cube-split generator of this package:

* a cubic lattice of spacing h covers the bounding box;
* a cell is kept when its centre lies inside the body (the "centre-in" rule);
* every kept cell is split into 6 tetrahedra around its main diagonal
  (Kuhn split), which is conforming across cells; every tet is positively
  oriented, det[p2-p1, p3-p1, p4-p1] > 0, as scadyn requires
  (src/common.f90:466-469, src/projection.f90:140-143);
* vertices whose 8 surrounding cells are all kept are jittered uniformly by
  +-jitter*h per axis (seeded), boundary vertices stay on the lattice.

Node ids are lexicographic in (z, y, x) of the lattice; tets follow their cells
in (z, y, x) order, 6 per cell.  Arrays are returned 0-based:
coord (Nnode, 3) float64, etopol (N, 4) int32.

Bodies: sphere, Gaussian random ellipsoid (log-normal radial factor with a
seeded real spherical-harmonic series, the model of other/scadyn_mesh.py:439-593),
and a two-layer permittivity helper.
"""
import math

import numpy as np

# Kuhn split of the unit cell: path 0 -> e_a -> e_a+e_b -> (1,1,1) for each
# axis permutation (a, b, c).  Corners are indexed by bit (x=1, y=2, z=4).
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _kuhn_tets():
    tets = []
    for a, b, c in _PERMS:
        v0 = 0
        v1 = v0 | (1 << a)
        v2 = v1 | (1 << b)
        v3 = 7
        tets.append([v0, v1, v2, v3])
    out = []
    for t in tets:
        p = np.array([[(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1] for v in t], float)
        d = np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]], axis=1))
        if d < 0:
            t = [t[0], t[2], t[1], t[3]]
        out.append(t)
    return np.array(out, np.int64)


KUHN = _kuhn_tets()


def cube_split(inside, lo, hi, n_cells, jitter=0.1, seed=0):
    """Mesh the cells of an n_cells^3 lattice on [lo,hi]^3 whose centre is inside.

    inside(points (M,3)) -> bool (M,).
    """
    h = (hi - lo) / n_cells
    c = lo + h * (np.arange(n_cells) + 0.5)
    cz, cy, cx = np.meshgrid(c, c, c, indexing='ij')
    centres = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    keep = inside(centres).reshape(n_cells, n_cells, n_cells)      # [z][y][x]
    nv = n_cells + 1
    # vertex usage and interior test (all 8 surrounding cells kept)
    pad = np.zeros((n_cells + 2,) * 3, bool)
    pad[1:-1, 1:-1, 1:-1] = keep
    cnt = np.zeros((nv, nv, nv), np.int64)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                cnt += pad[dz:dz + nv, dy:dy + nv, dx:dx + nv]
    used = cnt > 0
    interior = cnt == 8
    vid = -np.ones((nv, nv, nv), np.int64)
    vid[used] = np.arange(int(used.sum()))
    iz, iy, ix = np.nonzero(used)
    coord = np.stack([lo + h * ix, lo + h * iy, lo + h * iz], axis=1).astype(np.float64)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter * h, jitter * h, size=coord.shape)
    inter = interior[iz, iy, ix]
    coord[inter] += jit[inter]
    kz, ky, kx = np.nonzero(keep)
    corner = np.zeros((kz.size, 8), np.int64)
    for bit in range(8):
        dx, dy, dz = bit & 1, (bit >> 1) & 1, (bit >> 2) & 1
        corner[:, bit] = vid[kz + dz, ky + dy, kx + dx]
    etopol = corner[:, KUHN].reshape(-1, 4).astype(np.int32)
    return coord, etopol


def sphere(n_cells=24, radius=1.0, jitter=0.1, seed=0):
    """Sphere of the given radius on an n_cells^3 lattice over [-r, r]^3."""
    def inside(p):
        return np.sum(p * p, axis=1) <= radius * radius
    return cube_split(inside, -radius, radius, n_cells, jitter, seed)


def ball(n_cells=8, radius=1.0, seed=0, jitter=0.0):
    """Sphere with a faceted (not staircase) surface: the whole cube-split
    lattice of [-r, r]^3, every vertex moved radially by |p|_inf / |p|_2 so that
    the cube's concentric shells become spheres.  The map is a positive radial
    scaling (Jacobian determinant s^3 > 0 with s = |p|_inf / |p|_2), so the
    tets keep their orientation; boundary vertices lie on the sphere and the
    surface error is that of flat facets, O(h^2)."""
    c, e = cube_split(lambda p: np.ones(len(p), bool), -radius, radius, n_cells, jitter, seed)
    r2 = np.sqrt(np.sum(c * c, axis=1))
    s = np.where(r2 > 0, np.max(np.abs(c), axis=1) / np.maximum(r2, 1e-300), 1.0)
    c = c * s[:, None]
    p = c[e]
    det = np.einsum('ni,ni->n', p[:, 3] - p[:, 0], np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]))
    if (det <= 0).any():
        raise ValueError('ball: a mapped tet lost its orientation')
    return c, e


def _real_sh_series(lmax, coeffs, theta, phi):
    """s(theta,phi) = sum_{l,m} P_l^m(cos t) (a_lm cos m phi + b_lm sin m phi)."""
    from scipy.special import lpmv
    s = np.zeros_like(theta)
    ct = np.cos(theta)
    for (l, m), (a, b) in coeffs.items():
        s += lpmv(m, l, ct) * (a * np.cos(m * phi) + b * np.sin(m * phi))
    return s


def gaussian_ellipsoid_radius(axes=(1.0, 0.8, 0.6), sigma=0.1, corr_angle_deg=20.0,
                              lmax=10, seed=0):
    """Radius function r(theta, phi) of a Gaussian random ellipsoid.

    r = r_ellipsoid(theta, phi) * exp(s - beta^2/2), beta = sqrt(ln(1+sigma^2)),
    s a real spherical-harmonic series with the modified-Gaussian covariance of
    correlation angle gamma (Legendre coefficients
    c_l ~ (2l+1) i_l(1/ell^2) exp(-1/ell^2), ell = 2 sin(gamma/2)),
    normalised so that Var(s) = beta^2.
    """
    from scipy.special import spherical_in
    beta = math.sqrt(math.log(1.0 + sigma * sigma))
    gamma = math.radians(corr_angle_deg)
    ell = 2.0 * math.sin(0.5 * gamma)
    ls = np.arange(0, lmax + 1)
    cl = (2 * ls + 1) * spherical_in(ls, 1.0 / ell ** 2) * np.exp(-1.0 / ell ** 2)
    cl[0] = 0.0
    cl = cl / cl.sum()
    rng = np.random.default_rng(seed)
    coeffs = {}
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            # variance of P_l^m term so that sum_m has variance cl[l]*beta^2 on average
            norm = (2.0 - (m == 0)) * math.factorial(l - m) / math.factorial(l + m)
            std = beta * math.sqrt(cl[l] * norm)
            a = rng.normal() * std
            b = rng.normal() * std if m > 0 else 0.0
            coeffs[(l, m)] = (a, b)
    ax = np.asarray(axes, float)

    def radius(theta, phi):
        u = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi),
                      np.cos(theta)], axis=-1)
        r_ell = 1.0 / np.sqrt(np.sum((u / ax) ** 2, axis=-1))
        s = _real_sh_series(lmax, coeffs, theta, phi)
        return r_ell * np.exp(s - 0.5 * beta * beta)
    return radius


def star_body(radius_fn, n_cells, extent, jitter=0.1, seed=0):
    """Mesh a star-shaped body r <= radius_fn(theta, phi) on [-extent, extent]^3."""
    def inside(p):
        r = np.sqrt(np.sum(p * p, axis=1))
        th = np.arccos(np.clip(p[:, 2] / np.maximum(r, 1e-300), -1, 1))
        ph = np.arctan2(p[:, 1], p[:, 0])
        return r <= radius_fn(th, ph)
    return cube_split(inside, -extent, extent, n_cells, jitter, seed)


def gaussian_ellipsoid(n_cells=53, seed=0, **kw):
    rfn = gaussian_ellipsoid_radius(seed=seed, **kw)
    th = np.linspace(0, np.pi, 91)
    ph = np.linspace(-np.pi, np.pi, 181)
    T, P = np.meshgrid(th, ph)
    ext = 1.02 * float(rfn(T, P).max())
    return star_body(rfn, n_cells, ext, jitter=0.1, seed=seed)


def tet_volumes(coord, etopol):
    c = coord[etopol]
    d = np.einsum('ni,ni->n', c[:, 3] - c[:, 0], np.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]))
    return np.abs(d) / 6.0


def rescale_to_radius(coord, etopol, a):
    """Scale so the volume-equivalent radius is `a` (src/h5io.f90:128-131)."""
    vol = 0.0
    for v in tet_volumes(coord, etopol):
        vol += v
    a_eff = (3.0 * vol / 4.0 / math.pi) ** (1.0 / 3.0)
    return coord * a / a_eff


def layered_param(coord, etopol, r_core, eps_core, eps_mantle, a_scale=1.0):
    """Per-tet permittivity: eps_core for centroid radius < r_core*a_scale."""
    c = coord[etopol].mean(axis=1)
    r = np.sqrt(np.sum(c * c, axis=1))
    return np.where(r < r_core * a_scale, complex(eps_core), complex(eps_mantle))
