"""Build libscadyn_amd.so (HIP, gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libscadyn_amd.so')
SOURCES = ['api.hip', 'geom.hip', 'setup.hip', 'fft.hip', 'fft_reg.hip', 'matvec.hip', 'krylov.hip', 'consumer.hip']
OBJDIR = os.path.join(HERE, '_build')

HIPCC_FLAGS = [
    '-O3', '-std=c++17', '--offload-arch=gfx950', '-fPIC',
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


HDF5_PREFIX = '/opt/conda'
H5_LIB = os.path.join(HERE, 'libscadyn_h5.so')


def build_h5(verbose=False):
    """libscadyn_h5.so: the HDF5 mesh / T-matrix file formats (host C, HDF5 C API),
    a library of its own so that the solver library does not depend on HDF5."""
    inc = os.path.join(HDF5_PREFIX, 'include')
    libdir = os.path.join(HDF5_PREFIX, 'lib')
    if not os.path.exists(os.path.join(inc, 'hdf5.h')):
        return None
    tmp = H5_LIB + '.tmp'
    cmd = ['gcc', '-O2', '-std=c11', '-D_DEFAULT_SOURCE', '-fPIC', '-shared', '-Wall', '-I' + inc, '-o', tmp,
           os.path.join(CSRC, 'h5io.c'), '-L' + libdir, '-lhdf5', '-Wl,-rpath,' + libdir]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('building libscadyn_h5.so failed')
    os.replace(tmp, H5_LIB)
    return H5_LIB


def build(verbose=False):
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    os.makedirs(OBJDIR, exist_ok=True)
    # one hipcc per translation unit, in parallel, then one link
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, src.replace('.hip', '.o'))
        cmd = [hipcc] + HIPCC_FLAGS + ['-c', '-o', obj, os.path.join(CSRC, src)]
        if verbose:
            print(' '.join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed.append(src)
    if failed:
        raise RuntimeError('hipcc failed on ' + ', '.join(failed))
    tmp = LIB + '.tmp'
    cmd = [hipcc, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('linking libscadyn_amd.so failed')
    os.replace(tmp, LIB)
    build_fortran(verbose)
    build_h5(verbose)
    return LIB


SELFTEST = os.path.join(OBJDIR, 'fft_selftest')
SELFTESTS = {'fft': ('fft_selftest.hip', SELFTEST), 'bmgs': ('bmgs_selftest.hip', os.path.join(OBJDIR, 'bmgs_selftest')),
             'gather': ('gather_selftest.hip', os.path.join(OBJDIR, 'gather_selftest')),
             'mgsb': ('mgsb_selftest.hip', os.path.join(OBJDIR, 'mgsb_selftest'))}


def build_selftest(verbose=False, which='fft'):
    """Host self-tests of kernel code (CPU only, AddressSanitizer on the host
    code): 'fft' the register FFT stages (links the library for the pass
    planner), 'bmgs' the block Gram-Schmidt bodies, 'gather' the LDS-staged
    per-cube gather."""
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    os.makedirs(OBJDIR, exist_ok=True)
    src, exe = SELFTESTS[which]
    cmd = [hipcc, '-O1', '-g', '-std=c++17', '--offload-arch=gfx950', '-Xarch_host', '-fsanitize=address',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-Wno-unused-command-line-argument',
           '-o', exe, os.path.join(CSRC, src), '-L' + HERE, '-lscadyn_amd', '-Wl,-rpath,' + HERE]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('building the ' + which + ' self-test failed')
    return exe


if __name__ == '__main__':
    print(build(verbose=True))
