"""ctypes access to oracle/_ref/libscadyn_ref.so (reference Fortran subset).

Test infrastructure only.  `load()` returns None when the library has not been
built (no reference tree or no amdflang), and the tests that need it skip.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, '_ref', 'libscadyn_ref.so')
_lib = None

dp = ctypes.POINTER(ctypes.c_double)
ip = ctypes.POINTER(ctypes.c_int32)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            return None
        # lazy binding: dsyev_/dgeev_ (LAPACK) are referenced by common.f90 routines
        # that this path never calls, and no LAPACK is installed
        L = ctypes.CDLL(LIB, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        L.rb_analytical.argtypes = [dp, dp, dp, dp]
        L.rb_int_v_g.argtypes = [dp, ctypes.c_double, dp, dp, ctypes.c_int, dp, dp]
        L.rb_int_s_g.argtypes = [dp, dp, dp, ctypes.c_double, dp, dp, ctypes.c_int, dp]
        L.rb_vv_dvdv.argtypes = [dp, dp, ctypes.c_double, dp, dp]
        L.rb_table.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ip]
        L.rb_bessel.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp]
        L.rb_legendre.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.rb_geometry.argtypes = [ctypes.c_int, ctypes.c_int, dp, ip, ip, dp, ip]
        for f in ('rb_analytical', 'rb_int_v_g', 'rb_int_s_g', 'rb_vv_dvdv', 'rb_table', 'rb_bessel',
                  'rb_legendre', 'rb_geometry'):
            getattr(L, f).restype = None
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _f(a):
    """Fortran (column-major) contiguous float64 copy."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def analytical(tri, n, r):
    out = np.zeros(2)
    load().rb_analytical(_d(_f(tri.T)), _d(_f(n)), _d(_f(r)), _d(out))
    return out


def int_V_G(rf, k, Pb, wb, Bc):
    out = np.zeros(2)
    load().rb_int_v_g(_d(_f(rf)), k, _d(_f(Pb.T)), _d(_f(wb)), len(wb), _d(_f(Bc.T)), _d(out))
    return out[0] + 1j * out[1]


def int_S_G(rf, tri, n, k, Pb, wb):
    out = np.zeros(2)
    load().rb_int_s_g(_d(_f(rf)), _d(_f(tri.T)), _d(_f(n)), k, _d(_f(Pb.T)), _d(_f(wb)), len(wb), _d(out))
    return out[0] + 1j * out[1]


def vv_dvdv(Tc, Bc, k):
    vv = np.zeros(2)
    dd = np.zeros(18)
    load().rb_vv_dvdv(_d(_f(Tc.T)), _d(_f(Bc.T)), k, _d(vv), _d(dd))
    D = (dd[0::2] + 1j * dd[1::2]).reshape(3, 3, order='F')
    return vv[0] + 1j * vv[1], D


def table(which, order):
    p = np.zeros(3 * 200)
    w = np.zeros(200)
    n = np.zeros(1, np.int32)
    load().rb_table(which, order, _d(p), _d(w), n.ctypes.data_as(ip))
    dim = 2 if which == 2 else 3
    return p[:dim * n[0]].reshape(n[0], dim), w[:n[0]]


def bessel_j(n, z):
    out = np.zeros(2 * (n + 1))
    load().rb_bessel(n, float(z.real), float(z.imag), _d(out))
    return out[0::2] + 1j * out[1::2]


def legendre(n, x):
    out = np.zeros(n + 1)
    load().rb_legendre(n, x, _d(out))
    return out


def geometry(coord, etopol0):
    N = etopol0.shape[0]
    ints = np.zeros(8, np.int32)
    dbls = np.zeros(8)
    tc = np.zeros(N, np.int32)
    et = np.asfortranarray((etopol0 + 1).T.astype(np.int32))
    load().rb_geometry(coord.shape[0], N, _d(_f(coord.T)), et.ctypes.data_as(ip), ints.ctypes.data_as(ip),
                       _d(dbls), tc.ctypes.data_as(ip))
    return {'Nx': ints[0], 'Ny': ints[1], 'Nz': ints[2], 'N_cubes': ints[3], 'N_tet_cube': ints[4],
            'delta': dbls[0], 'box_delta': dbls[1], 'min_coord': dbls[2:5].copy(), 'tet_cube': tc - 1}
