"""The HDF5 file formats at the boundary (libscadyn_h5.so) against an
independent reader, h5dump of the HDF5 distribution: dataset names, shapes,
types and values of the mesh file (src/h5io.f90:16-134,
other/scadyn_mesh.py:77-103) and of the T-matrix file (src/h5io.f90:429-558,
:562-678, :682-833), and round trips through the library's own readers."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from helpers import A_RADIUS

H5DUMP = shutil.which('h5dump') or '/opt/conda/bin/h5dump'
pytestmark = pytest.mark.skipif(not os.path.exists(H5DUMP), reason='h5dump is not installed')


def _header(path):
    out = subprocess.run([H5DUMP, '-H', path], capture_output=True, text=True, check=True).stdout
    ds = {}
    for m in re.finditer(r'DATASET "([^"]+)" \{\s*DATATYPE\s+(\S+)\s*DATASPACE\s+SIMPLE \{ \( ([^)]*) \)', out):
        ds[m.group(1)] = (m.group(2), tuple(int(v) for v in m.group(3).split(',')))
    return ds


def _values(path, name, tmp_path, dtype):
    """The raw little-endian values of one dataset, as h5dump writes them."""
    out = str(tmp_path / (name.replace('-', '_') + '.bin'))
    subprocess.run([H5DUMP, '-d', '/' + name, '-b', 'LE', '-o', out, path], capture_output=True, check=True)
    return np.fromfile(out, dtype=dtype)


def test_mesh_file_layout_and_round_trip(tmp_path):
    from scadyn_amd import h5io, meshgen
    c, e = meshgen.sphere(4, seed=0)
    param = meshgen.layered_param(c, e, 0.6, 3.5 + 0.2j, 2.0 + 0.02j)
    path = str(tmp_path / 'mesh.h5')
    h5io.write_mesh(path, c, e, param)
    hdr = _header(path)
    assert hdr['coord'] == ('H5T_IEEE_F64LE', (c.shape[0], 3))
    assert hdr['etopol'] == ('H5T_STD_I32LE', (e.shape[0], 4))
    assert hdr['param_r'] == ('H5T_IEEE_F64LE', (e.shape[0],))
    assert hdr['param_i'] == ('H5T_IEEE_F64LE', (e.shape[0],))
    assert np.array_equal(_values(path, 'etopol', tmp_path, '<i4').reshape(-1, 4), e + 1)   # 1-based
    assert np.array_equal(_values(path, 'coord', tmp_path, '<f8').reshape(-1, 3), c)
    assert np.array_equal(_values(path, 'param_i', tmp_path, '<f8'), param.imag)
    # read back, rescaled to the volume-equivalent radius (src/h5io.f90:128-131)
    c2, e2, p2 = h5io.read_mesh(path, A_RADIUS)
    assert np.array_equal(e2, e) and p2.shape == (1, e.shape[0]) and np.array_equal(p2[0], param)
    ref = meshgen.rescale_to_radius(c, e, A_RADIUS)
    assert np.max(np.abs(c2 - ref)) < 1e-15 * A_RADIUS * 10
    c3, _, _ = h5io.read_mesh(path)                       # a = 0: as stored
    assert np.array_equal(c3, c)


def test_T_file_layout_bar_slices_and_round_trip(tmp_path):
    from scadyn_amd import h5io
    rng = np.random.default_rng(4)
    nmaxs = [1, 3, 2]
    blocks = []
    for n in nmaxs:
        nm = (n + 1) ** 2 - 1
        blocks.append(tuple(rng.standard_normal((nm, nm)) + 1j * rng.standard_normal((nm, nm)) for _ in range(4)))
    wl = np.array([600e-9, 500e-9, 400e-9])
    ref_a = [2.5, 0.01, A_RADIUS]
    path = str(tmp_path / 'T.h5')
    # T_empty, then the bars in another order (singleT_write2file)
    h5io.T_empty(path, nmaxs, ref_a, wl)
    hdr = _header(path)
    ts = sum(((n + 1) ** 2 - 1) ** 2 for n in nmaxs)
    for nm in ('Taai_r', 'Taai_i', 'Tabi_r', 'Tabi_i', 'Tbai_r', 'Tbai_i', 'Tbbi_r', 'Tbbi_i'):
        assert hdr[nm] == ('H5T_IEEE_F64LE', (ts,))
        assert not _values(path, nm, tmp_path, '<f8').any()
    assert hdr['T-ref-a'] == ('H5T_IEEE_F64LE', (3,)) and hdr['T-wavlens'] == ('H5T_IEEE_F64LE', (3,))
    h5io.T_empty(path, nmaxs, ref_a, wl)                 # an existing file is kept
    for b in (3, 1, 2):
        h5io.write_T_bar(path, nmaxs, b, blocks[b - 1])
    # the flat layout read independently: bar blocks column-major, concatenated
    flat = _values(path, 'Tabi_r', tmp_path, '<f8') + 1j * _values(path, 'Tabi_i', tmp_path, '<f8')
    off = 0
    for n, blk in zip(nmaxs, blocks):
        nm = (n + 1) ** 2 - 1
        assert np.array_equal(flat[off:off + nm * nm].reshape((nm, nm), order='F'), blk[1])
        off += nm * nm
    assert np.array_equal(_values(path, 'T-wavlens', tmp_path, '<f8'), wl)
    assert np.array_equal(_values(path, 'T-ref-a', tmp_path, '<f8'), ref_a)
    got = h5io.read_T(path, nmaxs)
    for b in range(3):
        for q in range(4):
            assert np.array_equal(got['bars'][b][q], blocks[b][q])
    # the all-bars writer (T_write2file) gives the same file contents
    path2 = str(tmp_path / 'T2.h5')
    h5io.write_T(path2, nmaxs, blocks, ref_a, wl)
    got2 = h5io.read_T(path2)
    for q in range(4):
        assert np.array_equal(got2['flat'][q], got['flat'][q])


def test_errors_are_reported(tmp_path):
    from scadyn_amd import h5io
    with pytest.raises(h5io.H5Error, match='cannot open'):
        h5io.read_mesh(str(tmp_path / 'absent.h5'))
    path = str(tmp_path / 'T.h5')
    h5io.T_empty(path, [1], [1.0, 0.0, 1.0], [1.0])
    with pytest.raises(h5io.H5Error, match='does not hold'):
        h5io.write_T_bar(path, [2], 1, tuple(np.zeros((8, 8), complex) for _ in range(4)))


def test_library_exports_every_declared_symbol():
    from scadyn_amd import h5io
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    txt = re.sub(r'/\*.*?\*/', '', open(os.path.join(root, 'include', 'scadyn_h5.h')).read(), flags=re.S)
    names = sorted(set(re.findall(r'\b(scadyn_h5_[a-zA-Z_]+)\s*\(', txt)))
    assert len(names) == 10
    out = subprocess.run(['nm', '-D', '--defined-only', h5io.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(r'\bT ' + n + r'\b', out), n
