"""Request loop on the GPU: requests solved concurrently give the same T-matrix
as each request solved alone (same slot count, deterministic reductions):
tolerance 1e-13 relative."""
import numpy as np
import pytest

from helpers import rel, sphere_case

pytestmark = pytest.mark.gpu

CASES = [dict(n_cells=4, ka=2.0, eps=2.5 + 0.01j, nmax=2),
         dict(n_cells=5, ka=3.0, eps=3.0 + 0.1j, nmax=3),
         dict(n_cells=6, ka=1.5, eps=2.0 + 0.0j, nmax=2)]


def test_alone_equals_mix():
    from scadyn_amd import native
    inputs = []
    for cs in CASES:
        c, e, p, k = sphere_case(n_cells=cs['n_cells'], ka=cs['ka'], eps=cs['eps'])
        inputs.append((c, e, p, k, cs['nmax']))
    alone = []
    for c, e, p, k, nmax in inputs:
        ncol = 2 * ((nmax + 1) ** 2 - 1)
        print(f'[gpu-server] alone {e.shape[0]} tets nmax {nmax}', flush=True)
        op = native.Operator(c, e, p, k, batch=min(ncol, 32))
        alone.append(op.tmatrix(nmax)[0])
        op.close()
    print('[gpu-server] mix of 3 requests', flush=True)
    sv = native.Server(max_active=3)
    ids = [sv.submit(c, e, p, k, nmax) for c, e, p, k, nmax in inputs]
    results = [sv.wait(i) for i in ids]
    sv.close()
    for (T, it, times), Ta in zip(results, alone):
        assert times[0] <= times[1] <= times[2] <= times[3]
        for a, b in zip(T, Ta):
            if np.linalg.norm(b) > 0:
                assert rel(a, b) < 1e-13
