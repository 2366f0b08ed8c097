"""Build libpylians_hip.so (gfx950) in-tree with hipcc.

    python -m pylians3_amd.build            # compile every HIP source, link the library

Each .hip file under csrc/ is compiled to an object in parallel, then linked
against rocFFT.  Flags:
  -ffp-contract=off       keep the reference's float expression rounding (no FMA
                          contraction in the mass-assignment weights)
  -munsafe-fp-atomics     hardware float atomics (global_atomic_add_f32 and the
                          LDS float64 adds) instead of CAS loops
  -target-feature -packed-fp32-ops   no packed FP32 instructions
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libpylians_hip.so')
OBJDIR = os.path.join(HERE, '_build')
SOURCES = ['ctx.hip', 'ma.hip', 'pk.hip', 'plane.hip', 'fft.hip']
ARCH = os.environ.get('PYLIANS_HIP_ARCH', 'gfx950')

HIPCC = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
CFLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=%s' % ARCH, '-ffp-contract=off',
          '-munsafe-fp-atomics', '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-Wno-unused-result', '-I' + os.path.join(HERE, '..', 'include')]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + CFLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=%s' % ARCH, '-o', tmp] + objs + [
        '-L/opt/rocm/lib', '-lrocfft', '-lrocprofiler-sdk-roctx', '-Wl,-rpath,/opt/rocm/lib']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
