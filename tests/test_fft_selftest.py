"""The register line FFT's own stage code, run on the CPU (host build with
AddressSanitizer) against a direct long-double DFT: every (R1, R2) split the
pass planner can choose, zero-padded inputs, cropped outputs, partial column
tiles, forward / inverse / fused convolution, complex128 and complex64.

Tolerances (relative Frobenius over all outputs): 1e-14 complex128,
2e-6 complex64."""
import os
import subprocess

from scadyn_amd import build


def _stale(exe, deps):
    if not os.path.exists(exe):
        return True
    t = os.path.getmtime(exe)
    return any(os.path.getmtime(d) > t for d in deps)


def test_register_fft_stages_match_direct_dft():
    csrc = build.CSRC
    deps = [os.path.join(csrc, f) for f in ('fft_selftest.hip', 'fft_reg_core.hpp', 'fft.hpp', 'internal.hpp')]
    if not os.path.exists(build.LIB):
        build.build()
    deps.append(build.LIB)
    exe = build.SELFTEST
    if _stale(exe, deps):
        build.build_selftest()
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0')
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    tail = '\n'.join(r.stdout.splitlines()[-5:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert '0 failures' in tail
