"""Build libso3g_hip.so (the HIP kernels and their C ABI) for gfx950.

Each ``csrc/*.hip`` translation unit is compiled by ``hipcc`` (host-only
``csrc/*.cpp`` by ``g++``) in parallel and linked into
``so3g_amd/lib/libso3g_hip.so``.  The float64 pointing code
must reproduce the CPU path bit for bit, so FMA contraction is off
(``-ffp-contract=off``) and float32 denormals are kept.  Packed-FP32
instructions are disabled (``-packed-fp32-ops``).

Usage: ``python -m so3g_amd.build [--jobs N] [--force]``.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, 'lib', 'obj')
LIB = os.path.join(HERE, 'lib', 'libso3g_hip.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('SO3G_OFFLOAD_ARCH', 'gfx950')

CXX = os.environ.get('CXX', 'g++')
CXXFLAGS = ['-O3', '-std=c++17', '-fPIC', '-ffp-contract=off']
HIPFLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC',
    '-ffp-contract=off', '-fno-gpu-flush-denormals-to-zero',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-command-line-argument',
]


def _headers():
    hs = glob.glob(os.path.join(CSRC, '*.h'))
    hs.append(os.path.join(os.path.dirname(HERE), 'include', 'so3g_hip.h'))
    return hs


def _compile(src, force):
    base, ext = os.path.splitext(os.path.basename(src))
    obj = os.path.join(OBJ, base + ('.o' if ext == '.hip' else '_host.o'))
    newest_dep = max(os.path.getmtime(p) for p in [src] + _headers())
    if (not force) and os.path.exists(obj) and os.path.getmtime(obj) >= newest_dep:
        return obj, None
    if ext == '.hip':
        cmd = [HIPCC] + HIPFLAGS + ['-c', src, '-o', obj]
    else:   # host-only C++ (no device code)
        cmd = [CXX] + CXXFLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f'{" ".join(cmd)}\n{r.stdout}\n{r.stderr}'
    return obj, None


def build(jobs=None, force=True, verbose=False):
    """Compile every HIP source and link the shared library.  Returns the
    library path.  Raises RuntimeError with the compiler output on failure."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, '*.hip')) + glob.glob(os.path.join(CSRC, '*.cpp')))
    jobs = jobs or min(len(srcs), os.cpu_count() or 4)
    objs, errors = [], []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, force), srcs):
            objs.append(obj)
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError('hipcc failed:\n' + '\n'.join(errors))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB}')
    return LIB


if __name__ == '__main__':
    force = '--force' in sys.argv
    jobs = None
    if '--jobs' in sys.argv:
        jobs = int(sys.argv[sys.argv.index('--jobs') + 1])
    build(jobs=jobs, force=force, verbose=True)
