"""Host-code sanitizer runs (CPU only).

* The oracle's C code (oracle/ma_oracle.c, oracle/pk_oracle.c) built with
  AddressSanitizer + UndefinedBehaviorSanitizer into a standalone driver
  (tests/emul/oracle_san.c) that calls every exported routine on small grids,
  2D and 3D, all schemes, edge positions, and checks mass and mode totals.
* The tiled-MA index emulator (tests/emul/ma_index_emul.cpp, which includes the
  host-callable helpers of csrc/ma.hip) built with the same sanitizers on its
  host code only (-Xarch_host), over a reduced set of cases.
Any out-of-bounds access, use after free, signed overflow or other undefined
operation aborts the run.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ENV = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=1',
               UBSAN_OPTIONS='halt_on_error=1:print_stacktrace=1')


def test_oracle_under_asan_ubsan(tmp_path):
    gcc = shutil.which('gcc')
    if gcc is None:
        pytest.skip('gcc not found')
    exe = str(tmp_path / 'oracle_san')
    src = os.path.join(ROOT, 'tests', 'emul', 'oracle_san.c')
    cmd = [gcc, '-O1', '-g', '-fno-omit-frame-pointer', '-fsanitize=address,undefined',
           '-fno-sanitize-recover=all', '-ffp-contract=off', '-fopenmp', src, '-o', exe, '-lm']
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=SAN_ENV)
    assert r.returncode == 0, (r.stdout[-3000:] + r.stderr[-3000:])
    assert 'oracle sanitizer run: ok' in r.stdout


def _hipcc():
    for p in (shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if p and os.path.exists(p):
            return p
    return None


@pytest.mark.skipif(_hipcc() is None, reason='hipcc not found')
def test_ma_index_emulator_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / 'ma_index_emul_san')
    src = os.path.join(ROOT, 'tests', 'emul', 'ma_index_emul.cpp')
    csrc = os.path.join(ROOT, 'pylians3_amd', 'csrc')
    cmd = [_hipcc(), '-x', 'hip', '-O1', '-g', '-std=c++17', '--offload-arch=gfx950',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-Xarch_host', '-fsanitize=address', '-Xarch_host', '-fsanitize=undefined',
           '-Xarch_host', '-fno-sanitize-recover=all',
           '-DPYL_EMUL_QUICK=1', '-I', csrc, '-I', os.path.join(ROOT, 'include'),
           src, os.path.join(csrc, 'ctx.hip'), '-o', exe, '-lrocfft', '-lrocprofiler-sdk-roctx']
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=SAN_ENV)
    assert r.returncode == 0, (r.stdout[-3000:] + r.stderr[-3000:])
    assert 'all ok' in r.stdout
