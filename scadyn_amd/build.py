"""Build libscadyn_amd.so (HIP, gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libscadyn_amd.so')
SOURCES = ['api.hip', 'geom.hip', 'setup.hip', 'fft.hip', 'matvec.hip', 'krylov.hip']

HIPCC_FLAGS = [
    '-O3', '-std=c++17', '--offload-arch=gfx950', '-fPIC', '-shared',
    # packed FP32 instructions stay off in all device code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def build(verbose=False):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [hipcc] + HIPCC_FLAGS + ['-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed building libscadyn_amd.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
