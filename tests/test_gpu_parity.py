"""HIP path vs the oracle on a seeded 192-tet sphere (ka = 2, eps = 2.5+0.01i).

Tolerances (error budget: the 98x27 pseudo-inverse W+ of the projection has
cond(W) ~ 1.5e6, so two correct SVD codes give S* that differ by a few
cond*eps ~ 1e-9; everything downstream inherits that):
  S (FP64)                              5e-9 relative
  Fg (FP64 FFT)                         1e-12 relative
  near-zone blocks (complex64 storage)  2 ulp of complex64 per block norm
  matvec                                1e-8 relative
  solve (same iteration counts)         1e-8 relative
  T-matrix                              1e-7 relative (the parity bar is 1e-6)
"""
import numpy as np
import pytest

from helpers import oracle_operator, refined_sphere_case, rel, sphere_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f'[gpu-parity] {msg}', flush=True)


@pytest.fixture(scope='module')
def ops():
    from scadyn_amd import native
    c, e, param, k = sphere_case()
    say('create operator (192 tets, fft_pad=1)')
    g = native.Operator(c, e, param, k, fft_pad=1)
    o = oracle_operator()
    yield g, o
    g.close()


def test_geometry(ops):
    g, o = ops
    inf = g.info()
    assert (inf['Nx'], inf['Ny'], inf['Nz']) == (o.g.Nx, o.g.Ny, o.g.Nz)
    assert inf['N_cubes'] == o.g.N_cubes
    assert inf['N_tet_cube'] == o.g.N_tet_cube
    assert inf['n_blocks'] == o.sp_mat.shape[0]
    assert abs(inf['delta'] - o.g.delta) <= 1e-15 * o.g.delta


def test_projection_S(ops):
    g, o = ops
    say('export S')
    S = g.export('S')
    for ch, ref in enumerate((o.S, o.Sx, o.Sy, o.Sz)):
        assert rel(S[ch], ref) < 5e-9, ch


def test_green_fft(ops):
    g, o = ops
    say('export Fg')
    fg = g.export('fg')
    Nx, Ny, Nz = o.g.Nx, o.g.Ny, o.g.Nz
    ref = o.Fg[:Nz + 1, :Ny + 1, :Nx + 1] / (8.0 * Nx * Ny * Nz)
    assert rel(fg, ref) < 1e-12


def test_near_blocks(ops):
    g, o = ops
    say('export blocks')
    ind = g.export('blkind')
    assert np.array_equal(ind - 1, o.sp_ind)
    blk = g.export('blocks').astype(np.complex128)
    ref = o.sp_mat.astype(np.complex128)
    scale = np.abs(ref).reshape(len(ref), -1).max(axis=1)
    err = np.abs(blk - ref).reshape(len(ref), -1).max(axis=1) / scale
    assert err.max() < 2 * 2.0 ** -23, err.max()


def test_matvec(ops):
    from oracle import pfft_ref
    g, o = ops
    rng = np.random.default_rng(1)
    n = 3 * o.g.N
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    say('matvec x3')
    y = g.matvec(x)
    for r in range(3):
        assert rel(y[r], pfft_ref.matvec(o, x[r])) < 1e-8


def test_solve(ops):
    from oracle import pfft_ref
    g, o = ops
    rng = np.random.default_rng(2)
    n = 3 * o.g.N
    b = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    say('solve x2')
    x, it = g.solve(b)
    for r in range(2):
        xr, ite, _ = pfft_ref.gmres(lambda v: pfft_ref.matvec(o, v), b[r].copy())
        assert it[r] == ite
        assert rel(x[r], xr) < 1e-8


def test_vswf(ops):
    from oracle import pfft_ref
    g, o = ops
    c, e, param, k = sphere_case()
    say('vswf nmax=3')
    mat = g.vswf(3)
    ref = pfft_ref.vswf_matrix(c, e, k, 3)
    assert rel(mat, ref) < 1e-12


def test_tmatrix(ops):
    from oracle import pfft_ref
    g, o = ops
    c, e, param, k = sphere_case()
    say('tmatrix nmax=2')
    T, it = g.tmatrix(2)
    Tref, itref = pfft_ref.tmatrix(o, 2, c, e)
    for a, b in zip(T, Tref):
        assert rel(a, b) < 1e-7
    assert list(it) == list(itref)


def test_solve_cgs_dgks():
    # classical Gram-Schmidt + DGKS (opts.ortho = 1): same Krylov iterates up to rounding
    from oracle import pfft_ref
    from scadyn_amd import native
    c, e, param, k = sphere_case()
    say('solve x2 with ortho=1')
    g = native.Operator(c, e, param, k, fft_pad=1, ortho=1)
    o = oracle_operator()
    rng = np.random.default_rng(2)
    n = 3 * o.g.N
    b = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    x, it = g.solve(b)
    for r in range(2):
        xr, ite, _ = pfft_ref.gmres(lambda v: pfft_ref.matvec(o, v), b[r].copy())
        assert abs(int(it[r]) - ite) <= 1
        assert rel(x[r], xr) < 1e-7
    g.close()


@pytest.mark.skip(reason='block MGS is disabled on the device: its first GPU run faulted and the cause '
                         'is not found yet (the kernel code passes the host self-test under AddressSanitizer)')
@pytest.mark.parametrize('maxit', [20, 50])
def test_solve_block_mgs(maxit):
    """Block MGS (opts.ortho = 2: 8 MGS coefficients per pass via the in-pass Gram
    correction) against the one-pass-per-vector MGS on the device, at a tight
    tolerance so that the cycles run through several blocks and restarts.  The
    two differ only by rounding: iteration counts within one, solutions within
    1e-8 (both converged to 1e-10)."""
    from scadyn_amd import native
    c, e, param, k = sphere_case(n_cells=6)
    say(f'block MGS vs MGS, maxit {maxit}')
    a = native.Operator(c, e, param, k, tol=1e-10, maxit=maxit, restart=6)
    b = native.Operator(c, e, param, k, tol=1e-10, maxit=maxit, restart=6, ortho=2)
    rng = np.random.default_rng(31)
    n = 3 * e.shape[0]
    rhs = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    xa, ia = a.solve(rhs)
    xb, ib = b.solve(rhs)
    say(f'iterations {ia} vs {ib}')
    assert np.max(np.abs(ia.astype(int) - ib.astype(int))) <= 1
    assert max(rel(xb[r], xa[r]) for r in range(5)) < 1e-8
    a.close()
    b.close()


@pytest.mark.parametrize('bs', [1, 2, 4, 8])
def test_solve_mgs_blocked_vs_oracle(bs):
    """MGS with bs basis vectors per pass (opts.mgs_block) against the oracle's
    GMRES at the reference's settings: identical iteration counts, solutions
    within 1e-8 (the blocked coefficients equal MGS up to rounding)."""
    from oracle import pfft_ref
    from scadyn_amd import native
    c, e, param, k = sphere_case()
    say(f'solve x2 with mgs_block={bs}')
    g = native.Operator(c, e, param, k, fft_pad=1, mgs_block=bs)
    o = oracle_operator()
    rng = np.random.default_rng(2)
    n = 3 * o.g.N
    b = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    x, it = g.solve(b)
    for r in range(2):
        xr, ite, _ = pfft_ref.gmres(lambda v: pfft_ref.matvec(o, v), b[r].copy())
        assert it[r] == ite
        assert rel(x[r], xr) < 1e-8
    g.close()


@pytest.mark.parametrize('bs,maxit', [(2, 20), (4, 20), (4, 50), (8, 50)])
def test_solve_mgs_blocked_deep(bs, maxit):
    """Blocked MGS against one vector per pass on the device, at a tight
    tolerance so that the cycles run through many blocks and restarts, with 5
    systems in one batch (slots at different columns): iteration counts within
    one, solutions within 1e-8 (both converged to 1e-10)."""
    from scadyn_amd import native
    c, e, param, k = sphere_case(n_cells=6)
    say(f'mgs_block {bs} vs 1, maxit {maxit}')
    a = native.Operator(c, e, param, k, tol=1e-10, maxit=maxit, restart=6, mgs_block=1)
    b = native.Operator(c, e, param, k, tol=1e-10, maxit=maxit, restart=6, mgs_block=bs)
    rng = np.random.default_rng(31)
    n = 3 * e.shape[0]
    rhs = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    xa, ia = a.solve(rhs)
    xb, ib = b.solve(rhs)
    say(f'iterations {ia} vs {ib}')
    assert np.max(np.abs(ia.astype(int) - ib.astype(int))) <= 1
    assert max(rel(xb[r], xa[r]) for r in range(5)) < 1e-8
    a.close()
    b.close()


@pytest.mark.parametrize('nmax', [1, 3, 8])
def test_project_mfma_gemm(nmax):
    """mat^H X on the FP64 matrix cores vs numpy, at 2nm = 6, 30 and 160 columns
    (one partial 16x16 block up to several 64x64 tiles with a ragged edge).
    Tolerance 1e-13 relative: FP64 products, different summation order."""
    from scadyn_amd import native
    c, e, param, k = sphere_case()
    g = native.Operator(c, e, param, k)
    nm = (nmax + 1) ** 2 - 1
    rng = np.random.default_rng(20 + nmax)
    X = rng.standard_normal((2 * nm, 3 * e.shape[0])) + 1j * rng.standard_normal((2 * nm, 3 * e.shape[0]))
    say(f'project nmax {nmax} ({2 * nm} columns)')
    T = g.project(nmax, X)
    mat = g.vswf(nmax)                      # (3N, 2nm)
    ref = mat.conj().T @ X.T                # (2nm, 2nm)
    assert rel(T, ref) < 1e-13
    g.close()


def test_per_tet_permittivity():
    """Per-tet permittivity (two layers by centroid radius plus a seeded
    per-tet perturbation): eps enters the near-zone diagonal term
    delta_TB / (eps_T - 1) and the tets are renumbered by cube on the device,
    so a tet-order mistake shows here and not with a uniform eps.  Near-zone
    blocks, matvec and the nmax = 2 T-matrix against the oracle, at the
    tolerances of the uniform case."""
    from oracle import pfft_ref
    from scadyn_amd import native
    c, e, _, k = sphere_case()
    cen = c[e].mean(axis=1)
    r = np.linalg.norm(cen, axis=1)
    rng = np.random.default_rng(17)
    param = np.where(r < 0.6 * r.max(), 3.5 + 0.2j, 2.0 + 0.02j) + 0.05 * rng.standard_normal(len(r))
    say('layered eps: 192 tets')
    g = native.Operator(c, e, param, k, fft_pad=1)
    o = pfft_ref.build_operator(c, e, param, k)
    blk = g.export('blocks').astype(np.complex128)
    ref = o.sp_mat.astype(np.complex128)
    scale = np.abs(ref).reshape(len(ref), -1).max(axis=1)
    err = np.abs(blk - ref).reshape(len(ref), -1).max(axis=1) / scale
    assert err.max() < 2 * 2.0 ** -23, err.max()
    x = rng.standard_normal((2, 3 * o.g.N)) + 1j * rng.standard_normal((2, 3 * o.g.N))
    y = g.matvec(x)
    for q in range(2):
        assert rel(y[q], pfft_ref.matvec(o, x[q])) < 1e-8
    T, it = g.tmatrix(2)
    Tref, itref = pfft_ref.tmatrix(o, 2, c, e)
    for a, b in zip(T, Tref):
        assert rel(a, b) < 1e-7
    assert list(it) == list(itref)
    g.close()


def test_matvec_dense_cubes_vs_oracle():
    """Half-refined sphere, 480 tets, cubes of up to 14 tets: at batch 64 the
    per-cube gather runs two 8-row passes in the dense cubes.  Every slot's
    matvec against the oracle's, beside the slot-per-lane kernels' (24 at a
    time).  This operator's output cancels strongly (near-zone and far-field
    parts sum to about 66 times the result's norm), so the S* differences of
    the pseudo-inverse (2e-9, checked here) come out near 1.6e-7 in both: the
    per-cube path is held to the distance of the per-lane one (2x), and both
    to 1e-6."""
    from oracle import pfft_ref
    from scadyn_amd import native
    c, e, param, k = refined_sphere_case(n_cells=4)
    say('oracle operator (480 tets)')
    o = pfft_ref.build_operator(c, e, param, k)
    g = native.Operator(c, e, param, k, fft_pad=1)
    assert g.info()['N_tet_cube'] == o.g.N_tet_cube > 8
    rng = np.random.default_rng(5)
    n = 3 * o.g.N
    x = rng.standard_normal((64, n)) + 1j * rng.standard_normal((64, n))
    say('matvec x64 and 24 at a time')
    y = g.matvec(x)
    yn = np.concatenate([g.matvec(x[i:i + 24]) for i in range(0, 64, 24)])
    ref = [pfft_ref.matvec(o, x[r]) for r in range(64)]
    wide = max(rel(y[r], ref[r]) for r in range(64))
    narrow = max(rel(yn[r], ref[r]) for r in range(64))
    say(f'worst relative difference to the oracle: per-cube {wide:.2e}, per-lane {narrow:.2e}')
    S = g.export('S')
    srel = max(rel(S[ch], ref_s) for ch, ref_s in enumerate((o.S, o.Sx, o.Sy, o.Sz)))
    near = pfft_ref.near_matvec(o, x[0])
    cancel = (np.linalg.norm(near) + np.linalg.norm(ref[0] - near)) / np.linalg.norm(ref[0])
    say(f'S* relative difference {srel:.2e}; near + far over total norm {cancel:.1f}')
    assert srel < 5e-9
    assert wide < 1e-6 and narrow < 1e-6 and wide < 2 * narrow
    g.close()


def test_tmatrix_leading_block_independent_of_nmax():
    """The T-matrix of nmax 1 is the leading 3 x 3 part of each block of the
    nmax-2 one: every column is its own solve and every row its own
    projection, with index n(n+1)+m-1 and M-waves before N-waves (a wrong
    column or row offset of the N-waves breaks it)."""
    from scadyn_amd import native
    c, e, param, k = sphere_case()
    op = native.Operator(c, e, param, k, tol=1e-11, restart=40)
    T1, _ = op.tmatrix(1)
    T2, _ = op.tmatrix(2)
    op.close()
    worst = max(rel(b[:3, :3], a) for a, b in zip(T1, T2))
    say(f'nmax 1 vs leading block of nmax 2: {worst:.2e}')
    assert worst < 1e-9
