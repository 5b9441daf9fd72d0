"""Build oracle/_ref/libscadyn_ref.so from the reference Fortran, in place.

Compiles, unmodified and from where they lie in the reference tree, the five
modules of the JVIE path that need neither FFTW nor HDF5:
src/common.f90, src/integration_points.f90, src/singularity_subtraction.f90,
src/sfunctions.f90 and src/geometry.f90, plus the test binding
oracle/ref_bind.f90.  Flags follow the reference Makefile (Makefile:4) without
-march=native.  Outputs go only to oracle/_ref/.  The rest of the path
(possu.f90 includes fftw3.f03; h5io.f90 uses the HDF5 Fortran module; both
absent here) is not built.

Test infrastructure only; the product never loads this library.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, '_ref')
LIB = os.path.join(OUT, 'libscadyn_ref.so')
REF_SRC = '/root/reference/src'
MODULES = ['common', 'integration_points', 'singularity_subtraction', 'sfunctions', 'geometry']


def available():
    return os.path.isdir(REF_SRC) and shutil.which('amdflang') is not None


def build(verbose=False):
    if not available():
        return None
    os.makedirs(OUT, exist_ok=True)
    moddir = os.path.join(OUT, 'mod')
    os.makedirs(moddir, exist_ok=True)
    srcs = [os.path.join(REF_SRC, m + '.f90') for m in MODULES] + [os.path.join(HERE, 'ref_bind.f90')]
    cmd = ['amdflang', '-O3', '-ffast-math', '-funroll-loops', '-fPIC', '-shared',
           '-module-dir', moddir, '-o', LIB + '.tmp'] + srcs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-4000:])
        raise RuntimeError('building the reference subset failed')
    os.replace(LIB + '.tmp', LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
