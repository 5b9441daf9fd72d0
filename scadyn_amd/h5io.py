"""ctypes binding of libscadyn_h5.so (include/scadyn_h5.h): the reference's mesh
and T-matrix HDF5 files, and a host driver in the form of calc_T
(src/T_matrix.f90:18-82) that reads a mesh file, solves each wavelength bar on
the GPU and writes the T-matrix file bar by bar.

Arrays on the Python side: coord (Nnode, 3) metres, etopol (N, 4) 0-based,
param (bars, N) complex; T blocks as the solver returns them, (nm, nm) complex.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libscadyn_h5.so')
_lib = None


class H5Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is not built; run scadyn_amd/build.py (or __graft_entry__.build())')
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64 = ctypes.c_int32, ctypes.c_int64
    L.scadyn_h5_last_error.restype = ctypes.c_char_p
    L.scadyn_h5_mesh_sizes.argtypes = [ctypes.c_char_p, P, P, P]
    L.scadyn_h5_read_mesh.argtypes = [ctypes.c_char_p, ctypes.c_double, P, P, P]
    L.scadyn_h5_write_mesh.argtypes = [ctypes.c_char_p, i32, i32, P, P, P]
    L.scadyn_h5_T_size.argtypes = [i32, P]
    L.scadyn_h5_T_size.restype = i64
    L.scadyn_h5_write_T.argtypes = [ctypes.c_char_p, i32, P, P, P, P, P, P, P]
    L.scadyn_h5_T_empty.argtypes = [ctypes.c_char_p, i32, P, P, P]
    L.scadyn_h5_write_T_bar.argtypes = [ctypes.c_char_p, i32, P, i32, P, P, P, P]
    L.scadyn_h5_T_sizes.argtypes = [ctypes.c_char_p, P, P]
    L.scadyn_h5_read_T.argtypes = [ctypes.c_char_p, P, P, P, P, P, P]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise H5Error(lib().scadyn_h5_last_error().decode())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _b(path):
    return os.fsencode(path)


def read_mesh(path, a=0.0):
    """(coord (Nnode,3), etopol (N,4) 0-based, param (bars, N) complex); with
    a > 0 the mesh is scaled to volume-equivalent radius a."""
    nn, nt, nb = (np.zeros(1, np.int32) for _ in range(3))
    _check(lib().scadyn_h5_mesh_sizes(_b(path), _p(nn), _p(nt), _p(nb)))
    coord = np.zeros((int(nn[0]), 3))
    etopol = np.zeros((int(nt[0]), 4), np.int32)
    param = np.zeros((int(nb[0]), int(nt[0])), np.complex128)
    _check(lib().scadyn_h5_read_mesh(_b(path), float(a), _p(coord), _p(etopol), _p(param)))
    return coord, etopol - 1, param


def write_mesh(path, coord, etopol, param):
    coord = np.ascontiguousarray(coord, np.float64)
    etopol = np.ascontiguousarray(np.asarray(etopol) + 1, np.int32)
    param = np.ascontiguousarray(np.broadcast_to(np.asarray(param, np.complex128), (etopol.shape[0],)))
    _check(lib().scadyn_h5_write_mesh(_b(path), coord.shape[0], etopol.shape[0], _p(coord), _p(etopol), _p(param)))


def _nm(nmax):
    return (nmax + 1) ** 2 - 1


def flat_T(blocks_per_bar):
    """Four complex [T_size] arrays in the file's layout from per-bar block tuples."""
    return [np.concatenate([np.asarray(b[q]).ravel(order='F') for b in blocks_per_bar]) for q in range(4)]


def write_T(path, nmaxs, blocks_per_bar, ref_a, wavelengths):
    nmax = np.ascontiguousarray(nmaxs, np.int32)
    T = [np.ascontiguousarray(t, np.complex128) for t in flat_T(blocks_per_bar)]
    ra = np.ascontiguousarray(ref_a, np.float64)
    wl = np.ascontiguousarray(wavelengths, np.float64)
    _check(lib().scadyn_h5_write_T(_b(path), nmax.size, _p(nmax), *[_p(t) for t in T], _p(ra), _p(wl)))


def T_empty(path, nmaxs, ref_a, wavelengths):
    nmax = np.ascontiguousarray(nmaxs, np.int32)
    ra = np.ascontiguousarray(ref_a, np.float64)
    wl = np.ascontiguousarray(wavelengths, np.float64)
    _check(lib().scadyn_h5_T_empty(_b(path), nmax.size, _p(nmax), _p(ra), _p(wl)))


def write_T_bar(path, nmaxs, bar, blocks):
    """bar is 1-based, as matrices%whichbar; blocks = (Taa, Tab, Tba, Tbb) (nm, nm)."""
    nmax = np.ascontiguousarray(nmaxs, np.int32)
    T = [np.ascontiguousarray(np.asarray(b).ravel(order='F'), np.complex128) for b in blocks]
    _check(lib().scadyn_h5_write_T_bar(_b(path), nmax.size, _p(nmax), int(bar), *[_p(t) for t in T]))


def read_T(path, nmaxs=None):
    """Flat (Taa, Tab, Tba, Tbb) [T_size], T-ref-a, T-wavlens; with nmaxs also
    the per-bar block tuples."""
    ts = np.zeros(1, np.int64)
    nb = np.zeros(1, np.int32)
    _check(lib().scadyn_h5_T_sizes(_b(path), _p(ts), _p(nb)))
    T = [np.zeros(int(ts[0]), np.complex128) for _ in range(4)]
    ra = np.zeros(3)
    wl = np.zeros(max(int(nb[0]), 1))
    _check(lib().scadyn_h5_read_T(_b(path), *[_p(t) for t in T], _p(ra), _p(wl)))
    out = {'flat': T, 'ref_a': ra, 'wavelengths': wl[:int(nb[0])]}
    if nmaxs is not None:
        bars, off = [], 0
        for n in nmaxs:
            nm = _nm(n)
            bars.append(tuple(t[off:off + nm * nm].reshape((nm, nm), order='F') for t in T))
            off += nm * nm
        out['bars'] = bars
    return out


def truncation_order(ka):
    """Nmax of size parameter ka: floor(ka + 4 ka^(1/3)) for ka > 1, else 4
    (src/common.f90:724-735)."""
    return int(math.floor(ka + 4.0 * ka ** (1.0 / 3.0))) if ka > 1.0 else 4


def solve_bars(mesh_path, T_path, wavelengths, a, nmaxs=None, whichbar=0, **opts):
    """The per-bar loop of calc_T (src/T_matrix.f90:36-79) with the GPU solve:
    read and rescale the mesh (read_mesh), create the zero T file if absent
    (T_empty), and for each bar (or only `whichbar`, 1-based) solve on the
    device and write the bar's slice (singleT_write2file).  wavelengths in
    metres (k = 2 pi / lambda).  As in the reference, every bar takes the
    first permittivity column unless one bar is selected, which takes its own
    (:37-43), and nmax defaults to the truncation order of
    k max(Nx, Ny, Nz) delta / 2 on the mesh's FFT grid (src/setup.f90:56-60)."""
    from scadyn_amd import native
    coord, etopol, param = read_mesh(mesh_path, a)
    wavelengths = np.asarray(wavelengths, float)
    ks = 2.0 * np.pi / wavelengths
    nbar = len(ks)
    bars = [whichbar] if whichbar else list(range(1, nbar + 1))
    col = param[whichbar - 1] if (whichbar and param.shape[0] > 1) else param[0]
    results, extent = {}, None
    for b in bars:
        op = native.Operator(coord, etopol, col, ks[b - 1], **opts)
        try:
            if extent is None:  # the grid is the mesh's alone (the same for every bar)
                inf = op.info()
                extent = max(inf['Nx'], inf['Ny'], inf['Nz']) * inf['delta']
            if nmaxs is None:
                nmaxs = [truncation_order(k * extent / 2.0) for k in ks]
            results[b] = op.tmatrix(int(nmaxs[b - 1]))
        finally:
            op.close()
    nmaxs = [int(n) for n in nmaxs]
    T_empty(T_path, nmaxs, [col[0].real, col[0].imag, a], wavelengths)
    for b in bars:
        write_T_bar(T_path, nmaxs, b, results[b][0])
    return nmaxs, results
