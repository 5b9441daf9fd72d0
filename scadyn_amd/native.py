"""ctypes binding of libscadyn_amd.so (the C ABI in include/scadyn_amd.h).

The library is built in-tree by scadyn_amd/build.py.  There is no fallback:
if the shared library is missing or fails to load, every entry point raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libscadyn_amd.so')

OK = 0
ERR_ARG, ERR_ORIENTATION, ERR_GRID, ERR_SINGULAR, ERR_NOCONV, ERR_MEMORY, ERR_DEVICE, ERR_UNSUPPORTED = range(1, 9)
PREC_FP64, PREC_MIXED = 0, 1
BASIS_AUTO, BASIS_C64, BASIS_F16 = 0, 1, 2  # storage of the mixed-precision Krylov basis
EXPORT_S, EXPORT_BLOCKS, EXPORT_BLKIND, EXPORT_FG = 1, 2, 3, 4
KCLASS = ['scatter', 'fftx', 'ffty', 'fftz_conv', 'iffty', 'ifftx', 'gather_near', 'krylov', 'tgemm']

# every symbol include/scadyn_amd.h declares
EXPORTED = ['scadyn_opts_default', 'scadyn_last_error', 'scadyn_tmat_solve', 'scadyn_op_create',
            'scadyn_op_destroy', 'scadyn_op_info', 'scadyn_op_export', 'scadyn_op_matvec', 'scadyn_op_matvec_prec',
            'scadyn_op_solve', 'scadyn_op_vswf', 'scadyn_op_project', 'scadyn_op_project_noalloc',
            'scadyn_op_tmatrix', 'scadyn_op_tmatrix_cols',
            'scadyn_tjob_create', 'scadyn_tjob_advance', 'scadyn_tjob_result', 'scadyn_tjob_destroy',
            'scadyn_op_profile',
            'scadyn_op_timings', 'scadyn_op_timings_prec', 'scadyn_cross_sections', 'scadyn_force_torque_z',
            'scadyn_footprint', 'scadyn_server_create', 'scadyn_server_submit',
            'scadyn_server_status', 'scadyn_server_wait', 'scadyn_server_cancel', 'scadyn_server_destroy']
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


class ScadynError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'scadyn error {code}: {msg}')
        self.code = code


class CMesh(ctypes.Structure):
    _fields_ = [('n_node', ctypes.c_int32), ('n_tet', ctypes.c_int32),
                ('coord', ctypes.POINTER(ctypes.c_double)),
                ('etopol', ctypes.POINTER(ctypes.c_int32)),
                ('param', ctypes.POINTER(ctypes.c_double))]


class CServerOpts(ctypes.Structure):
    _fields_ = [('device', ctypes.c_int32), ('max_active', ctypes.c_int32), ('hbm_budget', ctypes.c_int64),
                ('host_only', ctypes.c_int32), ('reserved', ctypes.c_int32 * 7)]


class COpts(ctypes.Structure):
    _fields_ = [('tol', ctypes.c_double), ('maxit', ctypes.c_int32), ('restart', ctypes.c_int32),
                ('M_ex', ctypes.c_int32), ('near_zone', ctypes.c_int32), ('grid_size', ctypes.c_double),
                ('order', ctypes.c_int32), ('precision', ctypes.c_int32), ('batch', ctypes.c_int32),
                ('strict', ctypes.c_int32), ('device', ctypes.c_int32), ('fft_pad', ctypes.c_int32),
                ('refine_tol', ctypes.c_double), ('ortho', ctypes.c_int32), ('fft_kernel', ctypes.c_int32),
                ('mgs_block', ctypes.c_int32), ('gather_kernel', ctypes.c_int32), ('basis', ctypes.c_int32),
                ('reserved', ctypes.c_int32 * 3)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f'{LIB_PATH} is not built; run scadyn_amd/build.py (or __graft_entry__.build())')
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    L.scadyn_opts_default.argtypes = [ctypes.POINTER(COpts)]
    L.scadyn_opts_default.restype = None
    L.scadyn_last_error.argtypes = []
    L.scadyn_last_error.restype = ctypes.c_char_p
    L.scadyn_tmat_solve.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(COpts), ctypes.c_double,
                                    ctypes.c_int32, dp, dp, dp, dp, ip]
    L.scadyn_op_create.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(COpts), ctypes.c_double,
                                   ctypes.POINTER(P)]
    L.scadyn_op_destroy.argtypes = [P]
    L.scadyn_op_destroy.restype = None
    L.scadyn_op_info.argtypes = [P, ctypes.POINTER(ctypes.c_int64), dp]
    L.scadyn_op_export.argtypes = [P, ctypes.c_int32, P, ctypes.c_int64]
    L.scadyn_op_matvec.argtypes = [P, ctypes.c_int32, dp, dp]
    L.scadyn_op_matvec_prec.argtypes = [P, ctypes.c_int32, ctypes.c_int32, dp, dp]
    L.scadyn_op_solve.argtypes = [P, ctypes.c_int32, dp, dp, ip]
    L.scadyn_op_vswf.argtypes = [P, ctypes.c_int32, dp]
    L.scadyn_op_project.argtypes = [P, ctypes.c_int32, dp, dp]
    L.scadyn_op_tmatrix.argtypes = [P, ctypes.c_int32, dp, dp, dp, dp, ip]
    L.scadyn_op_project_noalloc.argtypes = [P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, P, P, P]
    L.scadyn_op_tmatrix_cols.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ip, dp, ip]
    L.scadyn_tjob_create.argtypes = [P, ctypes.c_int32, ctypes.c_int32, ip, ctypes.c_int32, ctypes.POINTER(P)]
    L.scadyn_tjob_advance.argtypes = [P, ctypes.c_int32, ip]
    L.scadyn_tjob_result.argtypes = [P, dp, ip]
    L.scadyn_tjob_destroy.argtypes = [P]
    L.scadyn_tjob_destroy.restype = None
    L.scadyn_op_profile.argtypes = [P, ctypes.c_int32]
    L.scadyn_op_timings.argtypes = [P, dp]
    L.scadyn_op_timings_prec.argtypes = [P, ctypes.c_int32, dp]
    L.scadyn_cross_sections.argtypes = [ctypes.c_int32, dp, dp, dp, dp, ctypes.c_double, dp]
    L.scadyn_force_torque_z.argtypes = [ctypes.c_int32, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, dp]
    L.scadyn_footprint.argtypes = [ctypes.POINTER(CMesh), ctypes.POINTER(COpts), ctypes.c_int32,
                                   ctypes.POINTER(ctypes.c_int64)]
    L.scadyn_server_create.argtypes = [ctypes.POINTER(CServerOpts), ctypes.POINTER(P)]
    L.scadyn_server_submit.argtypes = [P, ctypes.POINTER(CMesh), ctypes.POINTER(COpts), ctypes.c_double,
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
    L.scadyn_server_status.argtypes = [P, ctypes.c_int64, ip, dp]
    L.scadyn_server_wait.argtypes = [P, ctypes.c_int64, dp, dp, dp, dp, ip, dp]
    L.scadyn_server_cancel.argtypes = [P, ctypes.c_int64]
    L.scadyn_server_destroy.argtypes = [P]
    L.scadyn_server_destroy.restype = None
    for name in EXPORTED:
        if name not in ('scadyn_opts_default', 'scadyn_last_error', 'scadyn_op_destroy', 'scadyn_server_destroy',
                        'scadyn_tjob_destroy'):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise ScadynError(rc, lib().scadyn_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def default_opts(**kw):
    o = COpts()
    lib().scadyn_opts_default(ctypes.byref(o))
    fields = {f[0] for f in COpts._fields_}
    for key, val in kw.items():
        if key not in fields:
            raise ScadynError(ERR_ARG, f'unknown option {key!r}')
        setattr(o, key, val)
    return o


class MeshArrays:
    """Keeps the contiguous arrays alive while the C mesh struct points at them.

    coord (Nnode,3) metres, etopol (N,4) 0-based, param (N,) complex.
    """

    def __init__(self, coord, etopol, param):
        self.coord = np.ascontiguousarray(coord, dtype=np.float64)
        self.etopol = np.ascontiguousarray(np.asarray(etopol) + 1, dtype=np.int32)
        self.param = np.ascontiguousarray(np.asarray(param, dtype=np.complex128))
        self.c = CMesh(self.coord.shape[0], self.etopol.shape[0], _dp(self.coord), _ip(self.etopol),
                       self.param.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


class Operator:
    """Device-resident pFFT operator for one mesh and wavenumber."""

    def __init__(self, coord, etopol, param, k, **opts):
        self.mesh = MeshArrays(coord, etopol, param)
        self.opts = default_opts(**opts)
        self.k = float(k)
        h = ctypes.c_void_p()
        check(lib().scadyn_op_create(ctypes.byref(self.mesh.c), ctypes.byref(self.opts), self.k,
                                     ctypes.byref(h)))
        self.h = h
        self.N = self.mesh.etopol.shape[0]

    def close(self):
        if getattr(self, 'h', None):
            lib().scadyn_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = np.zeros(16, np.int64)
        d = np.zeros(4)
        check(lib().scadyn_op_info(self.h, i.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(d)))
        keys = ['N_tet', 'Nx', 'Ny', 'Nz', 'N_cubes', 'N_tet_cube', 'n_blocks', 'Lx', 'Ly', 'Lz',
                'batch', 'device_bytes', 'outer_rounds', 'cube_rows', 'fft_tiles', 'fft_tiles_f32']
        out = {k: int(i[j]) for j, k in enumerate(keys)}
        for k in ('fft_tiles', 'fft_tiles_f32'):  # x, y, z, inverse y, inverse x
            out[k] = [(out[k] >> (8 * q)) & 0xff for q in range(5)]
        out['delta'], out['box_delta'] = float(d[0]), float(d[1])
        out['reorth_fired'] = int(d[3])
        return out

    def export(self, what):
        inf = self.info()
        N = self.N
        if what == 'S':
            a = np.zeros((4, N, 27), np.complex128)
            wid = EXPORT_S
        elif what == 'blocks':
            a = np.zeros((inf['n_blocks'], 3, 3), np.complex64)  # [p][j][i] col-major -> transpose below
            wid = EXPORT_BLOCKS
        elif what == 'blkind':
            a = np.zeros((inf['n_blocks'], 2), np.int32)
            wid = EXPORT_BLKIND
        elif what == 'fg':
            a = np.zeros((inf['Lz'] // 2 + 1, inf['Ly'] // 2 + 1, inf['Lx'] // 2 + 1), np.complex128)
            wid = EXPORT_FG
        else:
            raise ValueError(what)
        check(lib().scadyn_op_export(self.h, wid, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))
        if what == 'blocks':
            a = np.ascontiguousarray(a.transpose(0, 2, 1))  # -> [p][i][j]
        return a

    def matvec(self, x, precision=PREC_FP64):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex128)
        y = np.zeros_like(x)
        check(lib().scadyn_op_matvec_prec(self.h, precision, x.shape[0], _dp(x.view(np.float64)),
                                          _dp(y.view(np.float64))))
        return y

    def solve(self, rhs):
        rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=np.complex128)
        x = np.zeros_like(rhs)
        it = np.zeros(rhs.shape[0], np.int32)
        check(lib().scadyn_op_solve(self.h, rhs.shape[0], _dp(rhs.view(np.float64)),
                                    _dp(x.view(np.float64)), _ip(it)))
        return x, it

    def vswf(self, nmax):
        nm = (nmax + 1) ** 2 - 1
        mat = np.zeros((2 * nm, 3 * self.N), np.complex128)  # column-major (3N, 2nm)
        check(lib().scadyn_op_vswf(self.h, nmax, _dp(mat.view(np.float64))))
        return mat.T

    def project(self, nmax, X):
        """mat^H X for X of shape (2nm, 3N) (one solution per row)."""
        nm = (nmax + 1) ** 2 - 1
        Xc = np.ascontiguousarray(np.asarray(X, dtype=np.complex128))
        if Xc.shape != (2 * nm, 3 * self.N):
            raise ScadynError(ERR_ARG, f'X must have shape {(2 * nm, 3 * self.N)}')
        T = np.zeros((2 * nm, 2 * nm), np.complex128)  # filled column-major
        check(lib().scadyn_op_project(self.h, nmax, _dp(Xc.view(np.float64)), _dp(T.view(np.float64))))
        return T.T.copy()

    def project_noalloc(self, n, na, nx, mat_ptr, X_ptr, T_ptr):
        """T = mat^H X on device pointers (row-major complex128), no allocation."""
        check(lib().scadyn_op_project_noalloc(self.h, int(n), int(na), int(nx), ctypes.c_void_p(mat_ptr),
                                              ctypes.c_void_p(X_ptr), ctypes.c_void_p(T_ptr)))

    def tmatrix(self, nmax):
        nm = (nmax + 1) ** 2 - 1
        T = [np.zeros((nm, nm), np.complex128) for _ in range(4)]  # filled column-major
        it = np.zeros(2 * nm, np.int32)
        check(lib().scadyn_op_tmatrix(self.h, nmax, *[_dp(t.view(np.float64)) for t in T], _ip(it)))
        return tuple(t.T.copy() for t in T), it

    def tmatrix_cols(self, nmax, cols):
        """T(:, cols) (2nm x len(cols)), solving only those columns' systems."""
        nm = (nmax + 1) ** 2 - 1
        sel = np.ascontiguousarray(np.asarray(cols), dtype=np.int32)
        T = np.zeros((sel.size, 2 * nm), np.complex128)  # filled column-major
        it = np.zeros(sel.size, np.int32)
        check(lib().scadyn_op_tmatrix_cols(self.h, nmax, sel.size, _ip(sel), _dp(T.view(np.float64)), _ip(it)))
        return T.T.copy(), it

    def profile(self, on=True):
        check(lib().scadyn_op_profile(self.h, 1 if on else 0))

    def timings(self, precision=None):
        """Per-class totals; precision 'fp64' / 'mixed' (or PREC_*) restricts the
        matvec classes to the launches of that matvec precision."""
        t = np.zeros((len(KCLASS), 3))
        if precision is None:
            check(lib().scadyn_op_timings(self.h, _dp(t)))
        else:
            p = {'fp64': PREC_FP64, 'mixed': PREC_MIXED}.get(precision, precision)
            check(lib().scadyn_op_timings_prec(self.h, p, _dp(t)))
        return {name: {'ms': t[i, 0], 'launches': int(t[i, 1]), 'bytes': t[i, 2]}
                for i, name in enumerate(KCLASS)}


class TJob:
    """Resumable T-matrix solve over a column list (scadyn_tjob_*)."""

    def __init__(self, op, nmax, cols, spread=0):
        self.op = op  # keeps the operator alive
        self.nmax = nmax
        self.cols = np.ascontiguousarray(np.asarray(cols), dtype=np.int32)
        h = ctypes.c_void_p()
        check(lib().scadyn_tjob_create(op.h, nmax, self.cols.size, _ip(self.cols), int(spread), ctypes.byref(h)))
        self.h = h

    def advance(self, want):
        nd = np.zeros(1, np.int32)
        check(lib().scadyn_tjob_advance(self.h, int(want), _ip(nd)))
        return int(nd[0])

    def result(self):
        nm = (self.nmax + 1) ** 2 - 1
        T = np.zeros((self.cols.size, 2 * nm), np.complex128)
        it = np.zeros(self.cols.size, np.int32)
        check(lib().scadyn_tjob_result(self.h, _dp(T.view(np.float64)), _ip(it)))
        return T.T.copy(), it

    def close(self):
        if getattr(self, 'h', None):
            lib().scadyn_tjob_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tmat_solve(coord, etopol, param, k, nmax, **opts):
    """One wavelength bar through the single-call boundary (scadyn_tmat_solve)."""
    m = MeshArrays(coord, etopol, param)
    o = default_opts(**opts)
    nm = (nmax + 1) ** 2 - 1
    T = [np.zeros((nm, nm), np.complex128) for _ in range(4)]
    it = np.zeros(2 * nm, np.int32)
    check(lib().scadyn_tmat_solve(ctypes.byref(m.c), ctypes.byref(o), float(k), nmax,
                                  *[_dp(t.view(np.float64)) for t in T], _ip(it)))
    return tuple(t.T.copy() for t in T), it


def _blocks(T):
    nm = T[0].shape[0]
    nmax = int(round((nm + 1) ** 0.5)) - 1
    cols = [np.asfortranarray(np.asarray(t, np.complex128)) for t in T]
    return nmax, cols


def cross_sections(T, k):
    """Orientation-averaged (C_ext, C_sca, C_abs, Tr) of T = (Taa, Tab, Tba, Tbb)."""
    nmax, cols = _blocks(T)
    out = np.zeros(4)
    check(lib().scadyn_cross_sections(nmax, *[c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for c in cols],
                                      float(k), _dp(out)))
    return out


def force_torque_z(T, k, E, a):
    """(F_z, T_z, Q_f_z, Q_t_z) for an x-polarised plane wave along +z."""
    nmax, cols = _blocks(T)
    out = np.zeros(4)
    check(lib().scadyn_force_torque_z(nmax, *[c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for c in cols],
                                      float(k), float(E), float(a), _dp(out)))
    return out


def footprint(coord, etopol, param, nmax, **opts):
    """Peak device bytes of one T-matrix request (host-side estimate)."""
    m = MeshArrays(coord, etopol, param)
    o = default_opts(**opts)
    b = ctypes.c_int64()
    check(lib().scadyn_footprint(ctypes.byref(m.c), ctypes.byref(o), nmax, ctypes.byref(b)))
    return b.value


class Server:
    """The library's request loop: submit T-matrix requests, wait for results."""

    def __init__(self, device=0, max_active=8, hbm_budget=0, host_only=False):
        so = CServerOpts(device, max_active, hbm_budget, 1 if host_only else 0)
        h = ctypes.c_void_p()
        check(lib().scadyn_server_create(ctypes.byref(so), ctypes.byref(h)))
        self.h = h
        self.nmax = {}

    def submit(self, coord, etopol, param, k, nmax, **opts):
        m = MeshArrays(coord, etopol, param)
        o = default_opts(**opts)
        rid = ctypes.c_int64()
        check(lib().scadyn_server_submit(self.h, ctypes.byref(m.c), ctypes.byref(o), float(k), nmax,
                                         ctypes.byref(rid)))
        self.nmax[rid.value] = nmax
        return rid.value

    def status(self, rid):
        st = np.zeros(1, np.int32)
        t = np.zeros(4)
        check(lib().scadyn_server_status(self.h, rid, _ip(st), _dp(t)))
        return int(st[0]), t

    def wait(self, rid):
        nmax = self.nmax.pop(rid)
        nm = (nmax + 1) ** 2 - 1
        T = [np.zeros((nm, nm), np.complex128) for _ in range(4)]
        it = np.zeros(2 * nm, np.int32)
        t = np.zeros(4)
        check(lib().scadyn_server_wait(self.h, rid, *[_dp(x.view(np.float64)) for x in T], _ip(it), _dp(t)))
        return tuple(x.T.copy() for x in T), it, t

    def cancel(self, rid):
        check(lib().scadyn_server_cancel(self.h, rid))

    def close(self):
        if getattr(self, 'h', None):
            lib().scadyn_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
