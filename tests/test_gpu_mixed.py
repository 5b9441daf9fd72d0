"""Mixed precision (FP32 matvec inside FP64 iterative refinement) vs the FP64 path.

Tolerances:
  FP32 matvec vs FP64 matvec         2e-6 relative (complex64 S*, grids, FFTs)
  mixed solve vs tight FP64 solve    1e-8 relative (refine_tol 1e-10)
  mixed T vs tight FP64 T            1e-6 relative (the parity bar)
"""
import numpy as np
import pytest

from helpers import rel, sphere_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f'[gpu-mixed] {msg}', flush=True)


@pytest.fixture(scope='module')
def case():
    c, e, param, k = sphere_case(n_cells=6)
    return c, e, param, k


def test_fp32_matvec(case):
    from scadyn_amd import native
    c, e, param, k = case
    say(f'operator {e.shape[0]} tets')
    op = native.Operator(c, e, param, k)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((4, 3 * e.shape[0])) + 1j * rng.standard_normal((4, 3 * e.shape[0]))
    y64 = op.matvec(x)
    y32 = op.matvec(x, precision=native.PREC_MIXED)
    for r in range(4):
        assert rel(y32[r], y64[r]) < 2e-6
    op.close()


def test_mixed_solve_and_tmatrix(case):
    from scadyn_amd import native
    c, e, param, k = case
    say('fp64 tight')
    op64 = native.Operator(c, e, param, k, tol=1e-11, restart=20)
    say("mixed, cgs")
    opc = native.Operator(c, e, param, k, precision=native.PREC_MIXED, tol=1e-5, refine_tol=1e-10, ortho=1)
    say('mixed')
    opm = native.Operator(c, e, param, k, precision=native.PREC_MIXED, tol=1e-5, refine_tol=1e-10)
    rng = np.random.default_rng(6)
    b = rng.standard_normal((3, 3 * e.shape[0])) + 1j * rng.standard_normal((3, 3 * e.shape[0]))
    x64, _ = op64.solve(b)
    xm, _ = opm.solve(b)
    xc, _ = opc.solve(b)
    for r in range(3):
        assert rel(xm[r], x64[r]) < 1e-8
        assert rel(xc[r], x64[r]) < 1e-8
    T64, _ = op64.tmatrix(3)
    Tm, _ = opm.tmatrix(3)
    num = np.sqrt(sum(np.linalg.norm(a - b) ** 2 for a, b in zip(Tm, T64)))
    den = np.sqrt(sum(np.linalg.norm(b) ** 2 for b in T64))
    say(f'T relative difference {num / den:.2e}')
    assert num / den < 1e-6
    op64.close()
    opm.close()
    opc.close()


def test_mixed_restart_batches(case):
    """20 systems restarting together (more than one packed FP64 restart batch of
    8) with short cycles (maxit 10): every system reaches the true-residual
    target; solutions within 1e-8 of tight FP64 ones."""
    from scadyn_amd import native
    c, e, param, k = case
    rng = np.random.default_rng(8)
    b = rng.standard_normal((20, 3 * e.shape[0])) + 1j * rng.standard_normal((20, 3 * e.shape[0]))
    say('fp64 tight, 20 systems')
    op64 = native.Operator(c, e, param, k, tol=1e-11, restart=20)
    x64, _ = op64.solve(b)
    say('mixed, maxit 10')
    opm = native.Operator(c, e, param, k, precision=native.PREC_MIXED, tol=1e-3, refine_tol=1e-10, maxit=10,
                          restart=40)
    xm, _ = opm.solve(b)
    cycles = opm.info()['outer_rounds']
    say(f'restart cycles {cycles}')
    assert cycles >= 2
    for r in range(20):
        assert rel(xm[r], x64[r]) < 1e-8
    op64.close()
    opm.close()


def test_mixed_desynchronised_restarts(case):
    """Systems of very different difficulty in one mixed-precision batch (right-
    hand sides scaled over 6 orders and half of them smooth), so that slots
    restart at different steps while others are in their Arnoldi steps: the
    batch matvec then runs with restarting slots inactive, and partly filled
    narrow FP64 restart batches and full-width ones alternate.  Every solution
    within 1e-8 of a tight FP64 solve."""
    from scadyn_amd import native
    c, e, param, k = case
    rng = np.random.default_rng(21)
    n = 3 * e.shape[0]
    b = rng.standard_normal((24, n)) + 1j * rng.standard_normal((24, n))
    b[::2] = np.cumsum(b[::2], axis=1) / np.sqrt(n)   # smoother right-hand sides converge sooner
    b *= np.logspace(-3, 3, 24)[:, None]
    say('fp64 tight, 24 systems')
    op64 = native.Operator(c, e, param, k, tol=1e-11, restart=20)
    x64, it64 = op64.solve(b)
    say('mixed, maxit 12')
    opm = native.Operator(c, e, param, k, precision=native.PREC_MIXED, tol=1e-2, refine_tol=1e-10, maxit=12,
                          restart=60)
    xm, itm = opm.solve(b)
    say(f'fp64 iterations {it64.min()}..{it64.max()}, mixed {itm.min()}..{itm.max()}')
    for r in range(24):
        assert rel(xm[r], x64[r]) < 1e-8
    op64.close()
    opm.close()


def test_mixed_basis_storage(case):
    """The mixed solve with its Krylov basis stored in binary16 and in complex64
    (scadyn_opts.basis): both within 1e-8 of a tight FP64 solve, since the FP64
    true residual at every restart decides convergence, and the binary16 basis
    takes at most 1.5x the matvecs of the complex64 one."""
    from scadyn_amd import native
    c, e, param, k = case
    rng = np.random.default_rng(12)
    b = rng.standard_normal((6, 3 * e.shape[0])) + 1j * rng.standard_normal((6, 3 * e.shape[0]))
    say('fp64 tight, 6 systems')
    op64 = native.Operator(c, e, param, k, tol=1e-11, restart=20)
    x64, _ = op64.solve(b)
    its = {}
    for name, basis in (('complex64', native.BASIS_C64), ('binary16', native.BASIS_F16)):
        opm = native.Operator(c, e, param, k, precision=native.PREC_MIXED, tol=1e-4, refine_tol=1e-10,
                              restart=60, basis=basis)
        xm, it = opm.solve(b)
        its[name] = int(it.sum())
        say(f'{name} basis: matvecs {it.tolist()}')
        for r in range(6):
            assert rel(xm[r], x64[r]) < 1e-8
        opm.close()
    assert its['binary16'] <= 1.5 * its['complex64']
    op64.close()
