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


FORTRAN_DIR = os.path.join(HERE, 'fortran')
FORTRAN_BIN = os.path.join(FORTRAN_DIR, '_bin')
FORTRAN_EXE = os.path.join(FORTRAN_BIN, 'solve_bar')


def build_fortran(verbose=False):
    """Fortran host driver linked against the library (needs amdflang)."""
    import shutil
    fc = shutil.which('amdflang')
    if fc is None:
        return None
    os.makedirs(FORTRAN_BIN, exist_ok=True)
    cmd = [fc, '-O2', '-module-dir', FORTRAN_BIN, '-o', FORTRAN_EXE + '.tmp',
           os.path.join(FORTRAN_DIR, 'scadyn_amd_iface.f90'), os.path.join(FORTRAN_DIR, 'solve_bar_main.f90'),
           '-L' + HERE, '-lscadyn_amd', '-Wl,-rpath,$ORIGIN/../..']
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('building the Fortran host driver failed')
    os.replace(FORTRAN_EXE + '.tmp', FORTRAN_EXE)
    return FORTRAN_EXE


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
    build_fortran(verbose)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
