"""Batch-width kernel variants agree: batches whose last 64-slot chunk holds at
least 32 slots take the wave-per-node-block scatter and the wave-per-cube
gather (operands reused across a cube's tets; idle lanes of a partial chunk
load slot B-1 and store nothing); other widths take the slot-per-lane kernels.
Both must give the same matvec (summation order differs: 1e-13 relative in
FP64, 2e-6 in FP32) and the same GMRES solutions with identical iteration
counts (inactive slots in a partly filled batch included).  The matvec batch
is the number of right-hand sides passed, so the slot-per-lane side passes
them 24 at a time."""
import numpy as np
import pytest

from helpers import refined_sphere_case, rel, sphere_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f'[gpu-batch64] {msg}', flush=True)


@pytest.fixture(scope='module')
def case():
    return sphere_case(n_cells=7)


@pytest.mark.parametrize('wide', [64, 128, 56, 40])
def test_matvec_wide_vs_narrow(case, wide):
    from scadyn_amd import native
    c, e, param, k = case
    say(f'{e.shape[0]} tets, batch {wide} vs 24 at a time')
    a = native.Operator(c, e, param, k, precision=native.PREC_MIXED)
    b = native.Operator(c, e, param, k, precision=native.PREC_MIXED)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((wide, 3 * e.shape[0])) + 1j * rng.standard_normal((wide, 3 * e.shape[0]))
    for prec, tol in ((native.PREC_FP64, 1e-13), (native.PREC_MIXED, 2e-6)):
        ya = a.matvec(x, precision=prec)
        yb = np.concatenate([b.matvec(x[i:i + 24], precision=prec) for i in range(0, wide, 24)])
        worst = max(rel(ya[r], yb[r]) for r in range(wide))
        say(f'precision {prec}: worst relative difference {worst:.2e}')
        assert worst < tol
    a.close()
    b.close()


@pytest.mark.parametrize('narrow', [35, 24])
def test_solve_partly_filled_batch(case, narrow):
    from scadyn_amd import native
    c, e, param, k = case
    a = native.Operator(c, e, param, k, batch=64)
    b = native.Operator(c, e, param, k, batch=narrow)
    rng = np.random.default_rng(12)
    rhs = rng.standard_normal((70, 3 * e.shape[0])) + 1j * rng.standard_normal((70, 3 * e.shape[0]))
    xa, ia = a.solve(rhs)
    xb, ib = b.solve(rhs)
    say(f'iterations {ia[:8]} ... vs {ib[:8]} ...')
    assert np.array_equal(ia, ib)
    assert max(rel(xa[r], xb[r]) for r in range(70)) < 1e-12
    a.close()
    b.close()


@pytest.mark.parametrize('wide', [64, 112])
def test_matvec_dense_cubes_wide_vs_narrow(wide):
    """Cubes of up to 17 tets: the wave-per-cube gather runs several 8-row passes
    per cube, the wave-per-node-block scatter sums more tets per node."""
    from scadyn_amd import native
    c, e, param, k = refined_sphere_case()
    a = native.Operator(c, e, param, k, precision=native.PREC_MIXED)
    b = native.Operator(c, e, param, k, precision=native.PREC_MIXED)
    info = a.info()
    say(f'{e.shape[0]} tets, up to {info["N_tet_cube"]} per cube, batch {wide} vs 24 at a time')
    assert info['N_tet_cube'] > 8 and info['cube_rows'] == 1
    rng = np.random.default_rng(14)
    x = rng.standard_normal((wide, 3 * e.shape[0])) + 1j * rng.standard_normal((wide, 3 * e.shape[0]))
    # This operator's output cancels strongly (rounding differences come out
    # about 100x the unit roundoff): FP64 within 1e-12; in FP32 each kernel is
    # held to its distance from the FP64 result, the slot-chunk one no further
    # than 4x the slot-per-lane one.
    ya = a.matvec(x, precision=native.PREC_FP64)
    yb = np.concatenate([b.matvec(x[i:i + 24], precision=native.PREC_FP64) for i in range(0, wide, 24)])
    worst = max(rel(ya[r], yb[r]) for r in range(wide))
    say(f'FP64: worst relative difference {worst:.2e}')
    assert worst < 1e-12
    fa = a.matvec(x, precision=native.PREC_MIXED)
    fb = np.concatenate([b.matvec(x[i:i + 24], precision=native.PREC_MIXED) for i in range(0, wide, 24)])
    ea = max(rel(fa[r], ya[r]) for r in range(wide))
    eb = max(rel(fb[r], ya[r]) for r in range(wide))
    say(f'FP32 vs FP64: slot-chunk kernels {ea:.2e}, slot-per-lane kernels {eb:.2e}')
    assert ea < 1e-3 and ea < 4 * eb
    a.close()
    b.close()


@pytest.mark.parametrize('mesh', ['sphere', 'refined'])
def test_matvec_gather_on_matrix_cores(mesh):
    """The complex64 gather with the near-zone product on the matrix cores
    (gather_kernel = 1) against the vector-FMA kernel, both held to their
    distance from the FP64 matvec (rounding in another order): the matrix-core
    one no further than 4x the vector one, on the 7-cell sphere and on the
    refined sphere (cubes of up to 17 tets, several row passes per cube), whose
    output cancels strongly.  Full and partial slot chunks.  A dropped or
    misplaced near-zone term moves its rows by far more than rounding does."""
    from scadyn_amd import native
    c, e, p, k = sphere_case(n_cells=7) if mesh == 'sphere' else refined_sphere_case()
    rng = np.random.default_rng(2)
    for B in (64, 40):
        x = rng.standard_normal((B, 3 * e.shape[0])) + 1j * rng.standard_normal((B, 3 * e.shape[0]))
        ys = []
        for gk in (0, 1):
            op = native.Operator(c, e, p, k, batch=B, gather_kernel=gk)
            if gk == 0:
                y64 = op.matvec(x, precision=native.PREC_FP64)
            ys.append(op.matvec(x, precision=native.PREC_MIXED))
            op.close()
        ev, em = (max(rel(y[r], y64[r]) for r in range(B)) for y in ys)
        say(f'{mesh}, B={B}: FP32 vs FP64, vector gather {ev:.2e}, matrix-core gather {em:.2e}')
        assert ev < (1e-5 if mesh == 'sphere' else 1e-3)
        assert em < 4 * ev
