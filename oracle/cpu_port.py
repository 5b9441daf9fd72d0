"""CPU baseline: a multi-threaded port of the reference's matvec (test and
measurement infrastructure only; bench.py's cpu_baseline leg and tests/ use it,
the product never does).

The reference's own CPU path is not buildable in this image (possu.f90 needs
FFTW's fftw3.f03, h5io.f90 the HDF5 Fortran module), so the baseline is this
port of compute_Ax2 (src/gmres_module.f90:57-176) as the reference runs it:
six scatter -> zero-pad -> FFT -> x FFT(G) -> IFFT -> crop -> gather
convolutions on the reference's 2N-padded grid (src/build_G.f90,
src/possu.f90:140-289), twelve gathers, Y2 - k^2 Y, and the near-zone block
product.  Scatter, gather and near zone are OpenMP C (cpu_port.c); the FFTs
are scipy.fft (pocketfft) with the same thread count, where the reference calls
FFTW.  The operator arrays (S*, near-zone blocks) are exported from the device
build, so setup is not part of the sample.
"""
import ctypes
import os
import subprocess
import time

import numpy as np

from . import pfft_ref

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'cpu_port.c')
LIB = os.path.join(HERE, '_build', 'libcpu_port.so')
_lib = None


def build(out=LIB):
    """gcc -O3 -march=native -fopenmp (the reference's optimisation flags,
    Makefile:4, with OpenMP), for the host that compiles it."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + '.tmp'
    subprocess.run(['gcc', '-O3', '-march=native', '-fopenmp', '-fPIC', '-shared', '-o', tmp, SRC], check=True)
    os.replace(tmp, out)
    return out


def lib():
    """The port compiled for this host (-march=native), once per process."""
    global _lib
    if _lib is None:
        import tempfile
        L = ctypes.CDLL(build(os.path.join(tempfile.mkdtemp(prefix='cpu_port_'), 'libcpu_port.so')))
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.port_threads.restype = ctypes.c_int
        L.port_scatter.argtypes = [i64, P, P, P, i64, i64, P]
        L.port_gather.argtypes = [i64, P, P, P, ctypes.c_double, ctypes.c_double, P, i64]
        L.port_near.argtypes = [i64, i64, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class PortOperator:
    """Reference-layout operator arrays and the port's matvec."""

    def __init__(self, coord, etopol, k, S4, blocks, blkind):
        self.g = pfft_ref.build_geometry(coord, etopol)
        g = self.g
        self.k = k
        self.N = g.N
        self.S = [np.ascontiguousarray(S4[c]) for c in range(4)]   # S, Sx, Sy, Sz (N, 27)
        # box nodes of every tet's cube (src/geometry.f90:185-219), vectorised
        cube = g.tet_cube
        cx, cy, cz = cube % g.ncx, (cube // g.ncx) % g.ncy, cube // (g.ncx * g.ncy)
        o = np.arange(3)
        oz, oy, ox = np.meshgrid(o, o, o, indexing='ij')
        ix = 2 * cx[:, None] + ox.ravel()[None]
        iy = 2 * cy[:, None] + oy.ravel()[None]
        iz = 2 * cz[:, None] + oz.ravel()[None]
        self.indS = np.ascontiguousarray((ix + g.Nx * iy + g.Nx * g.Ny * iz).astype(np.int32))
        self.Fg = pfft_ref.build_Fg(g, k)
        # blocks [P][3][3] (i, j) -> column-major complex64 records
        self.blk = np.ascontiguousarray(np.asarray(blocks, np.complex64).transpose(0, 2, 1))
        self.ind = np.ascontiguousarray(np.asarray(blkind, np.int64) - 1, dtype=np.int32)
        self.ng = g.Nx * g.Ny * g.Nz

    def _conv(self, grid, workers):
        import scipy.fft
        g = self.g
        A = np.zeros(self.Fg.shape, complex)
        A[:g.Nz, :g.Ny, :g.Nx] = grid.reshape(g.Nz, g.Ny, g.Nx)
        A = scipy.fft.ifftn(self.Fg * scipy.fft.fftn(A, workers=workers), workers=workers)
        return np.ascontiguousarray(A[:g.Nz, :g.Ny, :g.Nx]).ravel()

    def matvec(self, x, workers):
        L = lib()
        N = self.N
        x = np.ascontiguousarray(x, complex)
        Y = np.zeros((N, 3), complex)
        Y2 = np.zeros((N, 3), complex)
        grid = np.zeros(self.ng, complex)
        xv = x.view(np.float64)
        Yf = Y.reshape(-1).view(np.float64)    # flat interleaved views of the (N, 3) arrays
        Y2f = Y2.reshape(-1).view(np.float64)
        for c in range(3):        # Y_c = S^T conv(S x_c)
            L.port_scatter(N, _p(self.S[0]), _p(self.indS), _p(xv[2 * c:]), 3, self.ng, _p(grid))
            gr = self._conv(grid, workers)
            L.port_gather(N, _p(self.S[0]), _p(self.indS), _p(gr), 1.0, 0.0, _p(Yf[2 * c:]), 3)
        for c in range(3):        # Y2_d += D_d^T conv(D_c x_c)
            L.port_scatter(N, _p(self.S[1 + c]), _p(self.indS), _p(xv[2 * c:]), 3, self.ng, _p(grid))
            gr = self._conv(grid, workers)
            for d in range(3):
                L.port_gather(N, _p(self.S[1 + d]), _p(self.indS), _p(gr), 1.0, 0.0,
                              _p(Y2f[2 * d:]), 3)
        y = (Y2 - self.k ** 2 * Y).ravel()
        L.port_near(N, self.ind.shape[0], _p(self.blk), _p(self.ind), _p(xv), _p(y.view(np.float64)))
        return y


def cpu_baseline(op_gpu, coord, etopol, param, k, matvecs_per_column, budget_s=25.0):
    """Columns/s of the port on the host's cores: its matvec on the full
    operator (arrays exported from the device build), timed over a bounded
    sample and scaled by the measured matvecs per column (the Krylov
    arithmetic, about a sixth of the reference's solve time, is not timed)."""
    t0 = time.time()
    port = PortOperator(coord, etopol, k, op_gpu.export('S'), op_gpu.export('blocks'), op_gpu.export('blkind'))
    prep = time.time() - t0
    threads = lib().port_threads()
    rng = np.random.default_rng(0)
    x = rng.standard_normal(3 * port.N) + 1j * rng.standard_normal(3 * port.N)
    n = 0
    t0 = time.time()
    while True:
        x = port.matvec(x, threads)
        x /= np.linalg.norm(x)
        n += 1
        if time.time() - t0 > budget_s or n >= 3:
            break
    per_mv = (time.time() - t0) / n
    return {'value': 1.0 / (per_mv * matvecs_per_column), 'unit': 'columns/s', 'cores': threads, 'kind': 'port',
            'sample': (f'{n} matvecs of the multi-threaded port of compute_Ax2 (OpenMP C scatter/gather/near '
                       f'zone, scipy.fft pocketfft, reference 2N padding, 6 convolutions) on the full '
                       f'{port.N}-tet operator with {threads} threads, {per_mv:.2f} s each, x '
                       f'{matvecs_per_column:.1f} matvecs/column of the FP64 GPU leg; the reference '
                       f'Fortran is not buildable here (no FFTW or HDF5 Fortran headers); '
                       f'{prep:.0f} s operator export')}
