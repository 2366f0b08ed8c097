"""The Gadget reader against the reference's own reader (library/readsnap.py).

tests/golden/gadget/ holds small binary snapshots (format I and II, both byte
orders, one and several files, header masses); tests/golden/gadget_golden.npz
holds what readsnap.snapshot_header and readsnap.read_block returned for them
(written by tests/golden/make_gadget_golden.py).  CPU only.
"""
import os

import numpy as np
import pytest

from pylians3_amd import readgadget as RG

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, 'golden')
CASES = ['f1_le_1', 'f2_le_3', 'f1_be_2', 'f2_be_1']


@pytest.fixture(scope='module')
def gold():
    with np.load(os.path.join(GOLD, 'gadget_golden.npz')) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize('case', CASES)
def test_header_matches_reference_reader(gold, case):
    h = RG.header(os.path.join(GOLD, 'gadget', case))
    for k in ('npart', 'massarr', 'nall'):
        np.testing.assert_array_equal(np.asarray(getattr(h, k)), gold['%s/%s' % (case, k)])
    for k in ('time', 'redshift', 'filenum', 'boxsize', 'omega_m', 'omega_l', 'hubble', 'format'):
        assert getattr(h, k) == gold['%s/%s' % (case, k)][()], k


@pytest.mark.parametrize('case', CASES)
@pytest.mark.parametrize('block', ['POS ', 'VEL ', 'ID  ', 'MASS'])
@pytest.mark.parametrize('ptype', [1, 2, 4])
def test_blocks_match_reference_reader(gold, case, block, ptype):
    """Blocks of one type, all files, physical velocities (v * sqrt(a) when z != 0)."""
    got = RG.read_block(os.path.join(GOLD, 'gadget', case), block, [ptype])
    exp = gold['%s/%s/%d' % (case, block.strip(), ptype)]
    assert got.shape == exp.shape
    if block in ('POS ', 'ID  '):
        assert got.dtype == exp.dtype
        np.testing.assert_array_equal(got, exp)
    elif block == 'VEL ':
        # the reference multiplies the float32 block by the float sqrt(a) in place
        np.testing.assert_array_equal(got.astype(np.float32), exp.astype(np.float32))
    else:
        np.testing.assert_array_equal(got.astype(np.float64), exp.astype(np.float64))
