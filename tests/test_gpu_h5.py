"""A mesh file in, a T-matrix file out: the per-bar driver (h5io.solve_bars,
the loop of calc_T) on the GPU, checked against the same solves through the
single-call boundary."""
import numpy as np
import pytest

from helpers import A_RADIUS, rel

pytestmark = pytest.mark.gpu


def test_solve_bars_from_and_to_files(tmp_path):
    from scadyn_amd import h5io, meshgen, native
    c, e = meshgen.sphere(4, seed=0)
    mesh = str(tmp_path / 'mesh.h5')
    h5io.write_mesh(mesh, c, e, np.full(e.shape[0], 2.5 + 0.01j))
    wl = 2 * np.pi * A_RADIUS / np.array([1.0, 1.5])        # ka 1 and 1.5
    Tpath = str(tmp_path / 'T.h5')
    print(f'[gpu-h5] {e.shape[0]} tets, 2 bars through the file driver', flush=True)
    nmaxs, res = h5io.solve_bars(mesh, Tpath, wl, A_RADIUS)
    assert len(nmaxs) == 2 and nmaxs[0] <= nmaxs[1]
    got = h5io.read_T(Tpath, nmaxs)
    assert np.allclose(got['wavelengths'], wl) and got['ref_a'][2] == A_RADIUS
    cs, es, ps = h5io.read_mesh(mesh, A_RADIUS)
    for b, nmax in enumerate(nmaxs):
        T, _ = native.tmat_solve(cs, es, ps[0], 2 * np.pi / wl[b], nmax)
        for q in range(4):
            if np.linalg.norm(T[q]) > 0:
                assert rel(got['bars'][b][q], T[q]) < 1e-13
