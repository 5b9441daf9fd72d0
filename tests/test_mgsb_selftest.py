"""MGS evaluated several basis vectors per pass (mgsb_core.hpp): the kernels'
own per-thread and per-slot code run on the CPU by the host self-test
(AddressSanitizer on the host build) against plain one-vector-at-a-time MGS, for
block sizes 2, 4 and 8, every slot phase (Arnoldi at block boundaries,
restarting, idle, done) and batch widths 5 to 300; and that the matvec input
block of every slot state lies inside the input array (complex128 basis with
x: m+2 blocks; complex64 basis of the mixed-precision solve: m+1).
Tolerance: 1e-12 relative on every H entry, norm and orthogonalised w."""
import os
import subprocess

from scadyn_amd import build


def test_blocked_mgs_matches_plain_mgs():
    src, exe = build.SELFTESTS['mgsb']
    deps = [os.path.join(build.CSRC, f) for f in (src, 'mgsb_core.hpp', 'krylov.hpp', 'internal.hpp')]
    if not os.path.exists(build.LIB):
        build.build()
    if not os.path.exists(exe) or any(os.path.getmtime(d) > os.path.getmtime(exe) for d in deps + [build.LIB]):
        build.build_selftest(which='mgsb')
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0')
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert '0 failures' in r.stdout
