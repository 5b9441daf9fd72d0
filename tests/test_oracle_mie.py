"""Pins the Mie oracle: published Bohren-Huffman values, unitarity, Rayleigh limit."""
import numpy as np

from oracle import mie


def test_bohren_huffman_published_case():
    # Bohren & Huffman, Appendix A sample run: m = 1.55, radius 0.525, lambda 0.6328
    x = 2 * np.pi * 0.525 / 0.6328
    qext, qsca, qback = mie.efficiencies(x, 1.55 + 0j, 20)
    assert abs(qext - 3.1054) < 5e-5
    assert abs(qsca - 3.1054) < 5e-5
    assert abs(qback - 2.9253) < 5e-5


def test_unitarity_lossless():
    # real m: every T eigenvalue t satisfies Re t = -|t|^2
    taa, tbb = mie.mie_tmatrix(3.0, 2.25, 10)
    for t in np.concatenate([taa, tbb]):
        assert abs(t.real + abs(t) ** 2) < 1e-12


def test_rayleigh_limit():
    # small sphere: Tbb(n=1) = -a_1 -> i (2/3) x^3 (m^2-1)/(m^2+2)
    x, eps = 1e-3, 2.5 + 0.1j
    taa, tbb = mie.mie_tmatrix(x, eps, 2)
    expect = 2j / 3 * x ** 3 * (eps - 1) / (eps + 2)
    assert abs(tbb[0] - expect) / abs(expect) < 1e-5


def test_convention_matches_reference_layout():
    # Taa = -b_n and Tbb = -a_n repeated over m = -n..n (src/T_matrix.f90:156-165)
    a, b = mie.mie_ab(2.0, np.sqrt(2.5 + 0.01j), 3)
    taa, tbb = mie.mie_tmatrix(2.0, 2.5 + 0.01j, 3)
    assert taa.shape == (15,)
    assert np.allclose(taa[:3], -b[0]) and np.allclose(taa[3:8], -b[1]) and np.allclose(taa[8:], -b[2])
    assert np.allclose(tbb[:3], -a[0])
