"""The C ABI library loads, exports every declared symbol, and reports host-side
errors (the reference's `stop` conditions) as status codes -- all without a GPU."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import sphere_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'scadyn_amd.h')


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(scadyn_[a-z_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    from scadyn_amd import native
    L = native.lib()
    names = _declared()
    assert len(names) >= 13
    out = subprocess.run(['nm', '-D', '--defined-only', native.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r'\bT (scadyn_\w+)', out))
    for n in names:
        assert n in exported, n
        getattr(L, n)
    assert set(native.EXPORTED) == set(names)


def test_default_options_are_the_reference_defaults():
    from scadyn_amd import native
    o = native.default_opts()
    # src/common.f90:80-89
    assert (o.tol, o.maxit, o.restart, o.M_ex, o.near_zone, o.grid_size, o.order) == \
        (1e-5, 50, 4, 2, 1, 1.8, 0)


def _create(coord, etopol, param, k=2e7, **kw):
    from scadyn_amd import native
    return native.Operator(coord, etopol, param, k, **kw)


def test_inverted_tet_is_an_orientation_error():
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    e = e.copy()
    e[5, [1, 2]] = e[5, [2, 1]]
    with pytest.raises(native.ScadynError) as ex:
        _create(c, e, p, k)
    assert ex.value.code == native.ERR_ORIENTATION
    assert 'detJ' in str(ex.value) and 'tet 6' in str(ex.value)


def test_bad_connectivity_is_an_argument_error():
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    e = e.copy()
    e[0, 0] = c.shape[0] + 5
    with pytest.raises(native.ScadynError) as ex:
        _create(c, e, p, k)
    assert ex.value.code == native.ERR_ARG


@pytest.mark.parametrize('kw', [{'M_ex': 1}, {'order': 1}, {'near_zone': 2}])
def test_unsupported_options(kw):
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    with pytest.raises(native.ScadynError) as ex:
        _create(c, e, p, k, **kw)
    assert ex.value.code == native.ERR_UNSUPPORTED


def test_nonpositive_wavenumber():
    from scadyn_amd import native
    c, e, p, k = sphere_case()
    with pytest.raises(native.ScadynError) as ex:
        _create(c, e, p, -1.0)
    assert ex.value.code == native.ERR_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from scadyn_amd import native
    monkeypatch.setattr(native, '_lib', None)
    monkeypatch.setattr(native, 'LIB_PATH', str(tmp_path / 'absent.so'))
    with pytest.raises(ImportError):
        native.lib()
