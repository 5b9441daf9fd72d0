"""Shared test inputs: seeded tiny meshes in the reference's units."""
import functools

import numpy as np

from scadyn_amd import meshgen

A_RADIUS = 1e-7          # params.in `a` default (src/common.f90:86)


@functools.lru_cache(maxsize=None)
def sphere_case(n_cells=4, ka=2.0, eps=2.5 + 0.01j, seed=0):
    """Cube-split sphere rescaled to volume-equivalent radius a (src/h5io.f90:128-131)."""
    c, e = meshgen.sphere(n_cells, seed=seed)
    c = meshgen.rescale_to_radius(c, e, A_RADIUS)
    k = ka / A_RADIUS
    param = np.full(e.shape[0], complex(eps))
    return c, e, param, k


def refined_sphere_case(n_cells=5, ka=2.0, eps=2.5 + 0.01j, seed=0):
    """sphere_case with every tet of the x < 0 half split into 4 at its centroid
    (same orientation, a quarter of the volume each): the grid step follows the
    mean tet volume, so the refined half's cubes hold up to 17 tets, more than
    one 8-row gather pass."""
    c, e, _, k = sphere_case(n_cells, ka, eps, seed)
    cent = c[e].mean(axis=1)
    extra, tets = [], []
    for t, tet in enumerate(e):
        if cent[t, 0] >= 0:
            tets.append(tet)
            continue
        m = c.shape[0] + len(extra)
        extra.append(cent[t])
        a, b, cc, d = tet
        tets += [[m, b, cc, d], [a, m, cc, d], [a, b, m, d], [a, b, cc, m]]
    c2 = np.vstack([c, np.array(extra)])
    e2 = np.array(tets, dtype=e.dtype)
    return c2, e2, np.full(e2.shape[0], complex(eps)), k


@functools.lru_cache(maxsize=None)
def oracle_operator(n_cells=4, ka=2.0, eps=2.5 + 0.01j, seed=0):
    from oracle import pfft_ref
    c, e, param, k = sphere_case(n_cells, ka, eps, seed)
    return pfft_ref.build_operator(c, e, param, k)


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
