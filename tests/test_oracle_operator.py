"""Pins of the oracle's pFFT operator pieces on tiny inputs.

* Green convolution (build_G + pruned FFT + crop/scale) vs the direct O(n^2) sum.
* Projection S*: translation invariance between tets that differ by a cube shift,
  and the least-squares fit of the tet's far field on the 98 test points.
* Near zone: for every stored near pair the assembled operator equals the exact
  (singular-integral) block -- the grid part cancels by construction -- and
  A - A^T is exactly the antisymmetric part of the self blocks.
* GMRES vs numpy.linalg.solve.
"""
import numpy as np
import pytest

from helpers import oracle_operator, rel, sphere_case
from oracle import pfft_ref as R


def test_green_convolution_direct_sum():
    c, e, param, k = sphere_case()
    g = R.build_geometry(c, e)
    Fg = R.build_Fg(g, k)

    class Op:
        pass
    op = Op()
    op.g, op.Fg = g, Fg
    rng = np.random.default_rng(0)
    n = g.Nx * g.Ny * g.Nz
    q = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = R._conv(op, q)
    iz, iy, ix = np.meshgrid(np.arange(g.Nz), np.arange(g.Ny), np.arange(g.Nx), indexing='ij')
    pts = R.node_xyz(g, ix.ravel(), iy.ravel(), iz.ravel())
    D = np.sqrt(np.sum((pts[:, None, :] - pts[None, :, :]) ** 2, axis=-1))
    with np.errstate(divide='ignore', invalid='ignore'):
        Gm = np.where(D > 0, np.exp(1j * k * D) / (4 * np.pi * D), 0.0)
    assert rel(got, Gm @ q) < 1e-12


def test_projection_translation_invariance():
    # a tet and its copy shifted by whole cubes sit at the same place in their
    # boxes, so their projection coefficients agree
    tet = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 1.1, 0.1], [0.1, 0.3, 0.9]]) * 1e-8
    vol = R.tetra_volume(tet)
    delta = 1.8 / 2 * vol ** R.F32_THIRD
    shift = np.array([3, 1, 2]) * 2 * delta
    coord = np.concatenate([tet, tet + shift])
    etopol = np.array([[0, 1, 2, 3], [4, 5, 6, 7]])
    g = R.build_geometry(coord, etopol)
    assert g.tet_cube[0] != g.tet_cube[1]
    S, Sx, Sy, Sz, _ = R.projection_const(g, 2e7)
    for a in (S, Sx, Sy, Sz):
        assert rel(a[1], a[0]) < 1e-8


def test_projection_fit_residual():
    # sum_j S_j G(p, node_j) sqrt(vol) reproduces int_tet G(p, r) dV on the 98 test points
    op = oracle_operator()
    g, k = op.g, op.k
    for t in (0, 17, 101):
        ids, _ = R.box_nodes(g, g.tet_cube[t])
        nodes = R.node_coords_of_ids(g, ids)
        cp = nodes[0] + g.box_delta / 2.0
        test = 2.0 * g.box_delta * R.quad.CUBE98_P + cp
        Pb, wb = R.linmap_tet(g.tets[t], R.quad.TET1_P, R.quad.TET1_W)
        Q = np.sum(R.Gr(test[:, None, :], Pb[None], k) * wb, axis=1)
        W = R.Gr(test[:, None, :], nodes[None], k)
        approx = W @ op.S[t] * np.sqrt(g.vol[t])
        assert rel(approx, Q) < 1e-3


def test_near_pairs_equal_exact_blocks():
    op = oracle_operator()
    A = R.dense_matrix(op)
    g, k = op.g, op.k
    _, _, exact = R.near_zone_const(g, op.param, np.zeros_like(op.S), np.zeros_like(op.S),
                                    np.zeros_like(op.S), np.zeros_like(op.S), k, return_f64=True)
    # with zero projections the blocks are the exact singular-integral blocks
    err = 0.0
    scale = np.abs(exact).max()
    for p, (T, B) in enumerate(op.sp_ind):
        blk = A[3 * T:3 * T + 3, 3 * B:3 * B + 3]
        err = max(err, np.abs(blk - exact[p]).max())
        if T != B:
            blkT = A[3 * B:3 * B + 3, 3 * T:3 * T + 3]
            err = max(err, np.abs(blkT - exact[p].T).max())
    # complex64 storage of the blocks bounds the agreement
    assert err / scale < 3e-7


def test_symmetry_defect_is_self_blocks_only():
    op = oracle_operator()
    A = R.dense_matrix(op)
    D = A - A.T
    for p, (T, B) in enumerate(op.sp_ind):
        if T == B:
            blk = op.sp_mat[p].astype(complex)
            D[3 * T:3 * T + 3, 3 * T:3 * T + 3] -= blk - blk.T
    assert np.abs(D).max() / np.abs(A).max() < 1e-12


def test_gmres_against_dense_solve():
    rng = np.random.default_rng(3)
    n = 60
    A = np.eye(n) * 3 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, ite, nmv = R.gmres(lambda v: A @ v, b, tol=1e-13, maxit=50, restart=4)
    assert rel(x, np.linalg.solve(A, b)) < 1e-11
    # the reference's counter: inner iterations that did not meet tol
    assert 0 < ite < 200


def test_gmres_restart_semantics():
    # a hard system that needs restarts: the result must still be the solution
    rng = np.random.default_rng(4)
    n = 80
    A = np.diag(np.linspace(1.0, 40.0, n)) + 0.05 * rng.standard_normal((n, n))
    b = rng.standard_normal(n) + 0j
    x, ite, nmv = R.gmres(lambda v: A @ v, b, tol=1e-10, maxit=10, restart=40)
    assert rel(x, np.linalg.solve(A, b)) < 1e-8
    assert ite > 10   # more than one cycle was used
