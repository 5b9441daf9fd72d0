"""Consumers of the T-matrix (host C ABI): orientation-averaged cross sections
(tr_T, src/mie.f90:728-759) and the z radiation force / torque of an
x-polarised plane wave (forcetorque, src/forces.f90:81-257), pinned with the
Mie T-matrix of a sphere against the Bohren-Huffman closed forms."""
import numpy as np
import pytest

from helpers import A_RADIUS, rel


def _mie_T(x, eps, nmax):
    from oracle import mie
    taa, tbb = mie.mie_tmatrix(x, eps, nmax)
    nm = taa.size
    z = np.zeros((nm, nm), complex)
    return (np.diag(taa), z, z.copy(), np.diag(tbb))


@pytest.mark.parametrize('x,eps', [(2.0, 2.5 + 0.01j), (5.213, 1.55 ** 2), (0.7, 3.5 + 0.2j)])
def test_cross_sections_and_radiation_pressure_of_a_sphere(x, eps):
    from oracle import mie
    from scadyn_amd import native
    nmax = int(x + 4 * x ** (1 / 3) + 4)
    T = _mie_T(x, eps, nmax)
    k = x / A_RADIUS
    crs = native.cross_sections(T, k)
    m = np.sqrt(complex(eps))
    qext, qsca, _ = mie.efficiencies(x, m, nmax)
    area = np.pi * A_RADIUS ** 2
    assert abs(crs[0] / area - qext) < 1e-10 * qext
    assert abs(crs[1] / area - qsca) < 1e-10 * qsca
    assert abs(crs[2] - (crs[0] - crs[1])) < 1e-12 * crs[0]
    # radiation pressure efficiency Q_pr = Q_ext - g Q_sca (Bohren & Huffman 4.61-4.62)
    a, b = mie.mie_ab(x, m, nmax)
    n = np.arange(1, nmax)
    gq = 4 / x ** 2 * (np.sum(n * (n + 2) / (n + 1) * np.real(a[:-1] * np.conj(a[1:]) + b[:-1] * np.conj(b[1:])))
                       + np.sum((2 * np.arange(1, nmax + 1) + 1) / (np.arange(1, nmax + 1) * np.arange(2, nmax + 2))
                                * np.real(a * np.conj(b))))
    F, Tz, Qf, Qt = native.force_torque_z(T, k, 1e4, A_RADIUS)
    assert abs(Qf - (qext - gq)) < 1e-9 * qext
    assert F > 0
    assert abs(Qt) < 1e-12 * Qf            # no spin torque on a sphere in linear polarisation


def test_optical_theorem_for_a_lossless_discretised_scatterer():
    """Real permittivity: C_abs of the discretised T (oracle, tightly solved)
    vanishes to the solver and discretisation level relative to C_ext."""
    from helpers import sphere_case
    from oracle import pfft_ref
    from scadyn_amd import native
    c, e, p, k = sphere_case(n_cells=4, ka=1.0, eps=2.25 + 0j)
    op = pfft_ref.build_operator(c, e, p, k)
    T, _ = pfft_ref.tmatrix(op, 3, c, e, tol=1e-10, maxit=100, restart=10)
    crs = native.cross_sections(T, k)
    assert abs(crs[2]) < 0.03 * crs[0]
    lossy = native.cross_sections(_mie_T(1.0, 2.25 + 0.5j, 6), k)
    assert lossy[2] > 0.05 * lossy[0]      # a lossy sphere absorbs


def test_bad_arguments():
    from scadyn_amd import native
    T = _mie_T(1.0, 2.0, 2)
    with pytest.raises(native.ScadynError):
        native.force_torque_z(T, -1.0, 1.0, A_RADIUS)
