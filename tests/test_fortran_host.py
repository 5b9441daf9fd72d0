"""The Fortran host path: scadyn_amd_iface.f90 (ISO_C_BINDING) driving one
wavelength bar the way the reference's calc_T does after the overlay.

The Fortran host executable (solve_bar) is built by scadyn_amd/build.py with amdflang; the tests
skip when it is absent.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel, sphere_case

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   'scadyn_amd', 'fortran', '_bin', 'solve_bar')
pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason='Fortran driver not built (no amdflang)')


def _run(tmp_path, coord, etopol0, param, k, nmax):
    fin = tmp_path / 'in.bin'
    fout = tmp_path / 'out.bin'
    with open(fin, 'wb') as f:
        np.array([coord.shape[0], etopol0.shape[0], nmax], np.int32).tofile(f)
        np.array([k], np.float64).tofile(f)
        np.ascontiguousarray(coord, np.float64).tofile(f)            # coord(3, n_node)
        np.ascontiguousarray(etopol0 + 1, np.int32).tofile(f)        # etopol(4, n_tet), 1-based
        np.ascontiguousarray(param, np.complex128).tofile(f)
    print(f'[fortran-host] {etopol0.shape[0]} tets nmax {nmax}', flush=True)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    raw = open(fout, 'rb').read()
    nm = (nmax + 1) ** 2 - 1
    rc = int(np.frombuffer(raw[:4], np.int32)[0])
    iters = np.frombuffer(raw[4:4 + 8 * nm], np.int32)
    T = np.frombuffer(raw[4 + 8 * nm:], np.complex128).reshape(4, nm, nm)
    return rc, iters, [t.T for t in T], r.stdout


def test_orientation_error_reaches_fortran(tmp_path):
    c, e, p, k = sphere_case()
    e = e.copy()
    e[3, [1, 2]] = e[3, [2, 1]]
    rc, _, _, out = _run(tmp_path, c, e, p, k, 1)
    assert rc == 2
    assert 'detJ' in out


@pytest.mark.gpu
def test_fortran_host_tmatrix(tmp_path):
    from oracle import pfft_ref
    from helpers import oracle_operator
    c, e, p, k = sphere_case()
    rc, iters, T, out = _run(tmp_path, c, e, p, k, 1)
    assert rc == 0, out
    Tref, itref = pfft_ref.tmatrix(oracle_operator(), 1, c, e)
    for a, b in zip(T, Tref):
        if np.linalg.norm(b) > 0:
            assert rel(a, b) < 1e-7
    assert list(iters) == list(itref)
    # the consumer called from Fortran gives the library's cross sections of the same T
    from scadyn_amd import native
    line = [ln for ln in out.splitlines() if ln.strip().startswith('cross sections')][0]
    crs_f = np.array([float(v) for v in line.split()[2:5]])
    crs = native.cross_sections(T, k)[:3]
    assert np.allclose(crs_f, crs, rtol=1e-12, atol=0.0), (crs_f, crs)
