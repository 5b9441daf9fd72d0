"""The two line-FFT kernel families agree on the device: the two-stage register
kernel (automatic choice) against the LDS Stockham kernel (fft_kernel = 1),
through FFT(G) at setup and through the matvec in both precisions, on axis
lengths that exercise several register splits (fft_pad 0 and 1, two mesh
sizes).

Tolerances (relative): Fg and FP64 matvec 1e-13 (two FP64 FFT orderings);
FP32 matvec 2e-6 (two FP32 FFT orderings, each ~1e-7 from exact)."""
import numpy as np
import pytest

from helpers import rel, sphere_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f'[gpu-fft] {msg}', flush=True)


@pytest.mark.parametrize('n_cells,fft_pad', [(6, 0), (6, 1), (16, 0)])
def test_register_vs_stockham(n_cells, fft_pad):
    from scadyn_amd import native
    c, e, param, k = sphere_case(n_cells=n_cells)
    say(f'{e.shape[0]} tets, fft_pad {fft_pad}')
    a = native.Operator(c, e, param, k, fft_pad=fft_pad, precision=native.PREC_MIXED)
    b = native.Operator(c, e, param, k, fft_pad=fft_pad, precision=native.PREC_MIXED, fft_kernel=1)
    inf = a.info()
    say(f"grid {inf['Nx']}x{inf['Ny']}x{inf['Nz']} fft {inf['Lx']}x{inf['Ly']}x{inf['Lz']}")
    assert rel(a.export('fg'), b.export('fg')) < 1e-13
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5, 3 * e.shape[0])) + 1j * rng.standard_normal((5, 3 * e.shape[0]))
    ya, yb = a.matvec(x), b.matvec(x)
    for r in range(5):
        assert rel(ya[r], yb[r]) < 1e-13
    ya, yb = a.matvec(x, precision=native.PREC_MIXED), b.matvec(x, precision=native.PREC_MIXED)
    for r in range(5):
        assert rel(ya[r], yb[r]) < 2e-6
    a.close()
    b.close()


def test_tuned_tiles():
    """A batch whose grids are large enough to be tuned (64 columns): the tile
    widths chosen are powers of two the kernels support, and the tuned matvec
    still equals the Stockham one."""
    from scadyn_amd import native
    c, e, param, k = sphere_case(n_cells=16)
    a = native.Operator(c, e, param, k, precision=native.PREC_MIXED)
    b = native.Operator(c, e, param, k, precision=native.PREC_MIXED, fft_kernel=1)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((64, 3 * e.shape[0])) + 1j * rng.standard_normal((64, 3 * e.shape[0]))
    ya, yb = a.matvec(x), b.matvec(x)
    inf = a.info()
    say(f"{e.shape[0]} tets, batch {inf['batch']}, tiles fp64 {inf['fft_tiles']} fp32 {inf['fft_tiles_f32']}")
    for t in inf['fft_tiles'] + inf['fft_tiles_f32']:
        assert t in (1, 2, 4, 8, 16, 32)
    for r in range(64):
        assert rel(ya[r], yb[r]) < 1e-13
    ya, yb = a.matvec(x, precision=native.PREC_MIXED), b.matvec(x, precision=native.PREC_MIXED)
    for r in range(64):
        assert rel(ya[r], yb[r]) < 2e-6
    a.close()
    b.close()


def test_register_radix_18_20():
    """Axis lengths above 256 (the 824k-tet sphere: 132 nodes per axis, L = 270
    = 15 x 18) take the radix-18/20 register kernel; it agrees with the LDS
    Stockham kernel through FFT(G) and the matvec in both precisions."""
    from scadyn_amd import native
    c, e, param, k = sphere_case(n_cells=64)
    a = native.Operator(c, e, param, k, precision=native.PREC_MIXED)
    inf = a.info()
    say(f"{e.shape[0]} tets, fft {inf['Lx']}x{inf['Ly']}x{inf['Lz']}")
    assert max(inf['Lx'], inf['Ly'], inf['Lz']) > 256
    b = native.Operator(c, e, param, k, precision=native.PREC_MIXED, fft_kernel=1)
    assert rel(a.export('fg'), b.export('fg')) < 1e-13
    rng = np.random.default_rng(7)
    x = rng.standard_normal((3, 3 * e.shape[0])) + 1j * rng.standard_normal((3, 3 * e.shape[0]))
    ya, yb = a.matvec(x), b.matvec(x)
    for r in range(3):
        assert rel(ya[r], yb[r]) < 1e-13
    ya, yb = a.matvec(x, precision=native.PREC_MIXED), b.matvec(x, precision=native.PREC_MIXED)
    for r in range(3):
        assert rel(ya[r], yb[r]) < 2e-6
    a.close()
    b.close()
