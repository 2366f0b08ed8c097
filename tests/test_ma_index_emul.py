"""CPU check of the tiled MA index logic (tests/emul/ma_index_emul.cpp).

The emulator includes csrc/ma.hip and runs its host-callable helpers over grid
sizes 1..64 (3D) and 1..200 (2D), every scheme, uniform / lattice / edge
positions, with and without weights; it asserts every index is in bounds,
every node is written once, and the grid equals a direct float64 deposit.
Needs hipcc (host compile only); no GPU is touched.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _hipcc():
    for p in (shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if p and os.path.exists(p):
            return p
    return None


@pytest.mark.skipif(_hipcc() is None, reason='hipcc not found')
def test_tiled_index_logic(tmp_path):
    exe = str(tmp_path / 'ma_index_emul')
    src = os.path.join(ROOT, 'tests', 'emul', 'ma_index_emul.cpp')
    csrc = os.path.join(ROOT, 'pylians3_amd', 'csrc')
    cmd = [_hipcc(), '-x', 'hip', '-O2', '-std=c++17', '--offload-arch=gfx950',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-I', csrc, '-I', os.path.join(ROOT, 'include'),
           src, os.path.join(csrc, 'ctx.hip'), '-o', exe, '-lrocfft', '-lrocprofiler-sdk-roctx']
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    out = r.stdout
    assert r.returncode == 0, out[-4000:] + r.stderr[-2000:]
    assert 'all ok' in out
    assert out.count(": ok") >= 380
    assert "dealing KR 4" in out and "dealing KR 2" in out and "count buckets: ok" in out
    assert "fix64: ok" in out
    assert "block window: ok" in out
    checked = [l for l in out.splitlines() if l.startswith('uniform finalize cells checked')]
    assert checked and int(checked[0].split(':')[1]) > 0
