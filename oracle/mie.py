"""Oracle: Mie T-matrix of a homogeneous sphere (test infrastructure only).

Follows the convention of the reference's mie_T_matrix
(src/T_matrix.f90:131-168, with src/sfunctions.f90:538-565): with
Bohren-Huffman coefficients a_n, b_n,

    Taa(n) = -b_n,   Tbb(n) = -a_n,   Tab = Tba = 0,

for every (n, m) index n(n+1)+m.  Unlike the reference, which drops Im(m)
(src/T_matrix.f90:149), the relative refractive index here is the complex
m = sqrt(eps), which is what the volume-integral solver sees.

Pinning: tests/test_oracle_mie.py checks the published Bohren-Huffman test case
(m = 1.55, x = 5.213: Qext = Qsca = 3.1054, Qback = 2.9253), unitarity
Re(T) = -|T|^2 for real m, and the Rayleigh limit.
"""
import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _psi(n, z):
    j = spherical_jn(n, z)
    dj = spherical_jn(n, z, derivative=True)
    return z * j, j + z * dj


def _xi(n, x):
    h = spherical_jn(n, x) + 1j * spherical_yn(n, x)
    dh = spherical_jn(n, x, derivative=True) + 1j * spherical_yn(n, x, derivative=True)
    return x * h, h + x * dh


def mie_ab(x, m, nmax):
    """Bohren-Huffman a_n, b_n for n = 1..nmax (size parameter x, relative index m)."""
    n = np.arange(1, nmax + 1)
    mx = m * x
    px, dpx = _psi(n, x)
    pm, dpm = _psi(n, mx)
    xx, dxx = _xi(n, x)
    a = (m * pm * dpx - px * dpm) / (m * pm * dxx - xx * dpm)
    b = (pm * dpx - m * px * dpm) / (pm * dxx - m * xx * dpm)
    return a, b


def mie_tmatrix(x, eps, nmax):
    """Diagonal Mie T blocks (Taa, Tbb) of length nmax(nmax+2)."""
    m = np.sqrt(complex(eps))
    a, b = mie_ab(x, m, nmax)
    ns = np.concatenate([[n] * (2 * n + 1) for n in range(1, nmax + 1)])
    return -b[ns - 1], -a[ns - 1]


def efficiencies(x, m, nmax):
    a, b = mie_ab(x, m, nmax)
    n = np.arange(1, nmax + 1)
    qext = 2.0 / x ** 2 * np.sum((2 * n + 1) * np.real(a + b))
    qsca = 2.0 / x ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    qback = np.abs(np.sum((2 * n + 1) * (-1.0) ** n * (a - b))) ** 2 / x ** 2
    return qext, qsca, qback
