"""LDS-staged per-cube gather: the kernel's own per-lane steps run on the CPU by
the host self-test (AddressSanitizer on the host build) against a direct
evaluation of grid gather + Y2 - k^2 Y + near zone, on a synthetic cube lattice
with empty cubes, cubes with more tets than one staging pass, transposed
blocks, inactive slots and batch widths 64, 112 and 56.
Tolerances: 1e-13 complex128, 1e-5 complex64 (relative, all outputs)."""
import os
import subprocess

from scadyn_amd import build


def test_lds_gather_steps_match_direct_sum():
    src, exe = build.SELFTESTS['gather']
    deps = [os.path.join(build.CSRC, f) for f in (src, 'gather_core.hpp', 'internal.hpp')]
    if not os.path.exists(build.LIB):
        build.build()
    if not os.path.exists(exe) or any(os.path.getmtime(d) > os.path.getmtime(exe) for d in deps + [build.LIB]):
        build.build_selftest(which='gather')
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0')
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert '0 failures' in r.stdout
