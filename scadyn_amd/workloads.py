"""Synthetic workloads of the T-matrix solve: the seeded meshes, materials,
wavenumbers and expansion orders that the benchmark and the profiling driver
run.  Lengths are in metres with the volume-equivalent radius a = 100 nm."""
import math

import numpy as np

A_RADIUS = 1e-7


def make_workload(name):
    from scadyn_amd import meshgen
    a = A_RADIUS
    if name == 'ellipsoid200k':
        c, e = meshgen.gaussian_ellipsoid(n_cells=56, seed=0)
        eps, ka, nmax = 3.0 + 0.1j, 8.0, 12
        desc = ('Gaussian random ellipsoid 1:0.8:0.6, sigma 0.1, corr 20 deg, l<=10, seed 0; '
                'cube-split; eps 3.0+0.1i; ka 8; nmax 12; FP64')
    elif name == 'layered824k':
        c, e = meshgen.sphere(64, seed=0)
        ka, nmax = 15.0, 20
        c = meshgen.rescale_to_radius(c, e, a)
        param = meshgen.layered_param(c, e, 0.6, 3.5 + 0.2j, 2.0 + 0.02j, a_scale=a)
        desc = ('sphere 64^3 cells, 2-layer eps (3.5+0.2i for r<0.6a, 2.0+0.02i), ka 15, nmax 20')
        return c, e, param, ka / a, nmax, desc
    elif name == 'sphere43k':
        c, e = meshgen.sphere(24, seed=0)
        eps, ka, nmax = 2.5 + 0.01j, 2.0, 5
        desc = 'sphere, 24^3 cells (centre-in), jitter 0.1h seed 0; eps 2.5+0.01i; ka 2; nmax 5; FP64'
    else:
        raise SystemExit(f'unknown workload {name}')
    c = meshgen.rescale_to_radius(c, e, a)
    param = np.full(e.shape[0], eps)
    return c, e, param, ka / a, nmax, desc


def nmax_for(x):
    """Expansion order of a wavelength bar: ceil(x + 4 x^(1/3) + 2)."""
    return int(math.ceil(x + 4.0 * x ** (1.0 / 3.0) + 2.0))


def ellipsoid_mesh(n_tets_target, seed):
    from scadyn_amd import meshgen
    n_cells = max(8, int(round(56 * (n_tets_target / 198390.0) ** (1.0 / 3.0))))
    c, e = meshgen.gaussian_ellipsoid(n_cells=n_cells, seed=seed)
    return meshgen.rescale_to_radius(c, e, A_RADIUS), e


def request_stream(name, rng):
    """(mesh, param, k, nmax, arrival offset s) lists for the request-loop workloads."""
    a = A_RADIUS
    reqs = []
    if name == 'bars16':
        # one Gaussian ellipsoid (~200k tets) at 16 size parameters log-spaced 2..15
        c, e = ellipsoid_mesh(198390, 0)
        for x in np.exp(np.linspace(np.log(2.0), np.log(15.0), 16)):
            reqs.append((c, e, np.full(e.shape[0], 3.0 + 0.1j), x / a, nmax_for(x), 0.0))
    elif name.startswith('serving'):
        # Poisson arrivals of random-ellipsoid requests; sizes log-uniform in [lo, hi] tets
        lo, hi, nreq, xmax = (20000, 400000, 24, 15.0) if name == 'serving' else (5000, 40000, 16, 6.0)
        t = 0.0
        for i in range(nreq):
            ntet = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
            x = float(rng.uniform(2.0, xmax))
            eps = complex(rng.uniform(1.5, 4.0), rng.uniform(0.0, 0.3))
            c, e = ellipsoid_mesh(ntet, seed=100 + i)
            reqs.append((c, e, np.full(e.shape[0], eps), x / a, nmax_for(x), t))
    return reqs
