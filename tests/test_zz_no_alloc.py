"""The no-allocation projection (scadyn_op_project_noalloc) against the
allocating one (scadyn_op_project) through the C ABI: two calls in a row, the
second smaller, on caller-held device buffers; sizes beyond the scratch limit
are rejected.  The caller's buffers come straight from the HIP runtime that
the library itself uses.  Bit-level agreement is not expected (the row order
of the sums differs): 1e-12 relative."""
import ctypes

import numpy as np
import pytest

from helpers import rel, sphere_case

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class DeviceBuffer:
    """A device allocation of the calling test (not of the library)."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), H2D) == 0

    def get(self, a):
        assert a.nbytes == self.nbytes and a.flags.c_contiguous
        assert self.hip.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(a.nbytes), D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _hip():
    from scadyn_amd import native
    native.lib()                                  # the library's HIP runtime is loaded first
    hip = ctypes.CDLL('libamdhip64.so.7')
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def test_project_noalloc_matches_project():
    from scadyn_amd import native
    c, e, p, k = sphere_case(n_cells=5)
    op = native.Operator(c, e, p, k)
    hip = _hip()
    rng = np.random.default_rng(11)
    n = 3 * e.shape[0]
    bufs = []
    for nmax in (4, 2):
        ncol = 2 * ((nmax + 1) ** 2 - 1)
        print(f'[no-alloc] {e.shape[0]} tets, nmax {nmax}: {ncol} x {ncol} projection', flush=True)
        mat = op.vswf(nmax)                                          # (3N, 2nm), reference rows
        X = rng.standard_normal((n, ncol)) + 1j * rng.standard_normal((n, ncol))
        T_ref = op.project(nmax, X.T)                                # mat^H X, allocating path
        mat_d, X_d = DeviceBuffer(hip, mat.nbytes), DeviceBuffer(hip, X.nbytes)
        T_d = DeviceBuffer(hip, ncol * ncol * 16)
        mat_d.put(mat)
        X_d.put(X)
        T_d.put(np.full((ncol, ncol), np.nan + 0j))
        op.project_noalloc(n, ncol, ncol, mat_d.p.value, X_d.p.value, T_d.p.value)
        T = T_d.get(np.empty((ncol, ncol), np.complex128))
        assert np.isfinite(T).all()
        assert rel(T, T_ref) < 1e-12
        bufs += [mat_d, X_d, T_d]
    with pytest.raises(native.ScadynError) as ex:
        op.project_noalloc(n, 4096, 2048, bufs[0].p.value, bufs[1].p.value, bufs[2].p.value)
    assert ex.value.code == native.ERR_ARG and 'no-allocation limit' in str(ex.value)
    for b in bufs:
        b.free()
    op.close()
