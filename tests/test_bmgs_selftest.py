"""Block modified Gram-Schmidt (opts.ortho = 2) kernel bodies, run on the CPU by
the host self-test (AddressSanitizer on the host build) against plain
one-vector-at-a-time MGS: every slot phase, Arnoldi columns at block edges,
batch widths 5..300 (one and two slot chunks), two Krylov dimensions.
Tolerance 1e-12 relative on the H column, |w|^2 before and after, and w."""
import os
import subprocess

from scadyn_amd import build


def test_block_mgs_bodies_match_mgs():
    src, exe = build.SELFTESTS['bmgs']
    deps = [os.path.join(build.CSRC, f) for f in (src, 'bmgs_core.hpp', 'krylov.hpp', 'internal.hpp')]
    if not os.path.exists(build.LIB):
        build.build()
    if not os.path.exists(exe) or any(os.path.getmtime(d) > os.path.getmtime(exe) for d in deps + [build.LIB]):
        build.build_selftest(which='bmgs')
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0')
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert '0 failures' in r.stdout
