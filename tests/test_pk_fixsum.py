"""CPU check of the power-spectrum binning's exact bin sums
(pylians3_amd/csrc/fixsum.h, run by tests/emul/pk_fixsum_emul.cpp): the split of a
double into fixed point (H * 2^48 + L) rounds x * 2^E to an integer, checked in
128-bit integers; sums of 10^5 contributions of a steep spectrum are the same bits
in any order and grouping and equal the 128-bit sum, and stay below 2^78; the low
word's addition limit leaves no overflow.  Needs hipcc (host
compile only); no GPU is touched.
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
def test_fixed_point_bin_sums(tmp_path):
    exe = str(tmp_path / 'pk_fixsum_emul')
    src = os.path.join(ROOT, 'tests', 'emul', 'pk_fixsum_emul.cpp')
    cmd = [_hipcc(), '-x', 'hip', '-O2', '-std=c++17', '--offload-arch=gfx950',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops', '-ffp-contract=off',
           src, '-o', exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    for line in ('split: ok', 'order independence: ok', 'bounds: ok', 'all ok'):
        assert line in r.stdout
