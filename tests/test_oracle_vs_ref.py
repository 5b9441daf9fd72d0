"""The oracle's restatement against the reference Fortran itself.

oracle/_ref/libscadyn_ref.so is compiled, unmodified, from the reference's
common/integration_points/singularity_subtraction/sfunctions/geometry modules
by oracle/build_ref.py (the rest of the path needs FFTW/HDF5 headers that are
not installed and is not built).  Skipped when that library is absent.
"""
import numpy as np
import pytest

from helpers import oracle_operator, sphere_case
from oracle import pfft_ref as R
from oracle import quadrature as Q
from oracle import ref_lib

pytestmark = pytest.mark.skipif(ref_lib.load() is None, reason='reference subset not built (oracle/build_ref.py)')

K = 2e7


def test_tables_equal_reference():
    p, w = ref_lib.table(1, 5)
    assert np.array_equal(p, Q.TET24_P) and np.array_equal(w, Q.TET24_W)
    p, w = ref_lib.table(1, 1)
    assert np.array_equal(p, Q.TET1_P) and np.array_equal(w, Q.TET1_W)
    p, w = ref_lib.table(2, 5)
    assert np.array_equal(p, Q.TRI7_P) and np.array_equal(w, Q.TRI7_W)
    p, w = ref_lib.table(2, 1)
    assert np.array_equal(p, Q.TRI1_P) and np.array_equal(w, Q.TRI1_W)
    p, _ = ref_lib.table(3, 0)
    assert {tuple(x) for x in p} == {tuple(x) for x in Q.CUBE98_P}


def _rand_tri(rng):
    return rng.standard_normal((3, 3)) * 1e-8


def test_analytical_integration_equals_reference():
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(200):
        tri = _rand_tri(rng)
        n = R.tri_normal(tri)
        r = tri.mean(axis=0) + rng.standard_normal(3) * 1e-8 * rng.choice([1e-3, 0.1, 1.0])
        a = np.array(R.analytical_integration(tri, n, r))
        b = ref_lib.analytical(tri, n, r)
        worst = max(worst, np.max(np.abs(a - b) / np.abs(b)))
    assert worst < 1e-12


def test_int_V_G_and_S_G_equal_reference():
    c, e, param, k = sphere_case()
    tets = c[e]
    rng = np.random.default_rng(1)
    for _ in range(40):
        t, b = rng.integers(0, len(tets), 2)
        Pt, _ = R.linmap_tet(tets[t], Q.TET24_P, Q.TET24_W)
        Pb, wb = R.linmap_tet(tets[b], Q.TET24_P, Q.TET24_W)
        rf = Pt[rng.integers(24)]
        a = R.int_V_G(rf[None], k, Pb, wb, tets[b])[0]
        ref = ref_lib.int_V_G(rf, k, Pb, wb, tets[b])
        assert abs(a - ref) / abs(ref) < 1e-12
        tri = tets[b][R.FACES[rng.integers(4)]]
        n = R.tri_normal(tri)
        Pb2, wb2 = R.linmap_tri(tri, Q.TRI7_P, Q.TRI7_W)
        a = R.int_S_G(rf[None], tri, n, k, Pb2, wb2)[0]
        ref = ref_lib.int_S_G(rf, tri, n, k, Pb2, wb2)
        assert abs(a - ref) / abs(ref) < 1e-12


def test_block_integrals_equal_reference():
    # V_V and dV_dV of near pairs, including self pairs and shared faces
    op = oracle_operator()
    tets = op.g.tets
    k = op.k
    pairs = op.sp_ind[::97]
    vv = R.integrate_V_V_G(tets[pairs[:, 0]], tets[pairs[:, 1]], k)
    dd = R.integrate_dV_dV_G(tets[pairs[:, 0]], tets[pairs[:, 1]], k)
    for i, (T, B) in enumerate(pairs):
        rv, rd = ref_lib.vv_dvdv(tets[T], tets[B], k)
        assert abs(vv[i] - rv) / abs(rv) < 1e-12
        assert np.abs(dd[i] - rd).max() / np.abs(rd).max() < 1e-12


def test_bessel_and_legendre_equal_reference():
    from scipy.special import spherical_jn, lpmv
    for x in (0.01, 0.2, 1.7, 6.0, 14.5):
        ref = ref_lib.bessel_j(25, complex(x))
        ours = spherical_jn(np.arange(26), x)
        assert np.allclose(ours, ref.real, rtol=1e-12, atol=0)
        assert np.abs(ref.imag).max() == 0.0
    for n in (1, 5, 13):
        for x in (-0.93, 0.1, 0.77):
            ref = ref_lib.legendre(n, x)
            ours = np.array([lpmv(m, n, x) for m in range(n + 1)])
            assert np.allclose(ours, ref, rtol=1e-12, atol=1e-300)


def test_reference_bessel_underflow_quirk():
    # cspherebessel seeds its backward recurrence with dcmplx(1.0d0 - 100) = -99
    # (src/sfunctions.f90:37) instead of 1e-100; for kr below ~1e-3 at orders
    # above ~20 the recurrence overflows and the reference returns zeros.  The
    # meshes here never put a tet centroid that close to the origin; the HIP VSWF
    # kernel (and the oracle, via scipy) compute the true j_n there.
    ref = ref_lib.bessel_j(25, complex(0.003))
    assert not np.any(np.abs(ref[:3]) > 0.5)


@pytest.mark.parametrize('n_cells', [4, 24])
def test_geometry_equals_reference(n_cells):
    c, e, param, k = sphere_case(n_cells=n_cells)
    g = R.build_geometry(c, e)
    ref = ref_lib.geometry(c, e)
    assert (g.Nx, g.Ny, g.Nz) == (ref['Nx'], ref['Ny'], ref['Nz'])
    assert g.N_cubes == ref['N_cubes'] and g.N_tet_cube == ref['N_tet_cube']
    assert g.delta == ref['delta'] and g.box_delta == ref['box_delta']
    assert np.array_equal(g.min_coord, ref['min_coord'])
    assert np.array_equal(g.tet_cube, ref['tet_cube'])
