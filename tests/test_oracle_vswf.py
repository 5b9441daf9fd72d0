"""Pins the oracle's regular VSWFs (src/mie.f90:408-496 convention).

* curl M = k N and curl N = k M (finite differences), which fixes alpha, beta,
  gamma, the vector harmonics and the spherical->cartesian map;
* the normalisation int |M_nm|^2 dOmega = j_n(kr)^2 on any sphere;
* the reference's phase convention: |m| in the Legendre function and no
  (-1)^m, hence M_{n,-m} = conj(M_{n,m}) and N_{n,-m} = conj(N_{n,m}) for real kr;
* the closed form M_10 = sqrt(3/(4 pi)) j_1(kr) sin(theta)/sqrt(2) phi_hat, from
  B_theta = q (P_2^0 - P_0^0) = -sin(theta)/sqrt(2) and C = (0, B_phi, -B_theta).
"""
import numpy as np
from scipy.special import spherical_jn

from oracle import pfft_ref as R

K = 2e7
NMAX = 3
R0 = np.array([0.31e-7, -0.22e-7, 0.47e-7])


def _fields(r):
    return R.calc_MN(NMAX, K, r)


def _curl(F, r0, h):
    J = None
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        d = (F(r0 + e) - F(r0 - e)) / (2 * h)
        if J is None:
            J = np.zeros((3, 3) + d.shape[1:], complex)
        J[:, a] = d
    return np.stack([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])


def test_curl_identities():
    M0, N0 = _fields(R0)
    h = 1e-4 / K
    cM = _curl(lambda r: _fields(r)[0], R0, h)
    cN = _curl(lambda r: _fields(r)[1], R0, h)
    assert np.abs(cM - K * N0).max() / np.abs(K * N0).max() < 1e-7
    assert np.abs(cN - K * M0).max() / np.abs(K * M0).max() < 1e-7


def test_normalisation_on_sphere():
    rr = 0.5e-7
    x, w = np.polynomial.legendre.leggauss(30)
    th = np.arccos(x)
    ph = np.linspace(0, 2 * np.pi, 60, endpoint=False)
    acc = 0.0
    for t, wt in zip(th, w):
        for p in ph:
            r = rr * np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])
            M, _ = _fields(r)
            acc = acc + np.sum(np.abs(M) ** 2, axis=0) * wt * (2 * np.pi / len(ph))
    n_of = np.concatenate([[n] * (2 * n + 1) for n in range(1, NMAX + 1)])
    assert np.allclose(acc, spherical_jn(n_of, K * rr) ** 2, rtol=1e-10)


def test_phase_convention():
    M, N = _fields(R0)
    for n in range(1, NMAX + 1):
        for m in range(1, n + 1):
            ip, im = n * (n + 1) + m - 1, n * (n + 1) - m - 1
            assert np.allclose(M[:, im], np.conj(M[:, ip]), rtol=1e-13, atol=0)
            assert np.allclose(N[:, im], np.conj(N[:, ip]), rtol=1e-13, atol=0)
        i0 = n * (n + 1) - 1
        assert np.abs(M[:, i0].imag).max() == 0.0


def test_M10_closed_form():
    M, _ = _fields(R0)
    r = np.linalg.norm(R0)
    th = np.arccos(R0[2] / r)
    ph = np.arctan2(R0[1], R0[0])
    phi_hat = np.array([-np.sin(ph), np.cos(ph), 0.0])
    # M_10 = cc j_1 (0, 0, -B_theta), B_theta = q (P_2^0 - P_0^0) = -sin(theta)/sqrt(2)
    expect = np.sqrt(3 / (4 * np.pi)) * spherical_jn(1, K * r) * np.sin(th) / np.sqrt(2) * phi_hat
    assert np.allclose(M[:, 1], expect, rtol=1e-12, atol=0)
