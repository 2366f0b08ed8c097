"""Batched MA / Pk / XPk: every realisation's result is bitwise the same as the
single call on it, and matches the oracle (uneven particle counts, empty
realisations, weights, two fields with XPk)."""
import numpy as np
import pytest

from conftest import assert_bins, assert_cells, contrib_counts

pytestmark = pytest.mark.gpu


def _realisations(nreal, dims, box, seed, weighted=True):
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, 30000, nreal)
    counts[1] = 0                                     # an empty realisation
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    pos = (rng.random((offs[-1], 3)) * box).astype(np.float32)
    W = rng.lognormal(0, 1, offs[-1]).astype(np.float32) if weighted else None
    return offs, pos, W


@pytest.mark.parametrize('mas', ['NGP', 'CIC', 'TSC', 'PCS'])
def test_ma_batch_equals_single_calls(oracle, mas):
    print('case ma_batch', mas, flush=True)
    from pylians3_amd import MAS_library as MASL
    from pylians3_amd.batch import MA_batch
    nreal, dims, box = 7, 24, 60.0
    offs, pos, W = _realisations(nreal, dims, box, 5)
    grids = np.zeros((nreal, dims, dims, dims), np.float32)
    MA_batch(pos, offs, grids, box, mas, W)
    for r in range(nreal):
        sl = slice(offs[r], offs[r + 1])
        single = np.zeros((dims,) * 3, np.float32)
        MASL.MA(pos[sl], single, box, mas, W[sl])
        np.testing.assert_array_equal(grids[r], single)
        ref = np.zeros((dims,) * 3, np.float32)
        oracle.ma_f32(pos[sl], ref, box, mas, W[sl])
        assert_cells(grids[r], ref, what='real %d' % r, nadds=contrib_counts(pos[sl], dims, box, mas))
    # plain atomic batch form agrees within tolerance
    g2 = np.zeros_like(grids)
    MA_batch(pos, offs, g2, box, mas, W, atomic=True)
    for r in range(nreal):
        assert_cells(g2[r], grids[r])


@pytest.mark.parametrize('axis', [0, 2])
def test_pk_batch_equals_single_calls(oracle, axis):
    print('case pk_batch', axis, flush=True)
    from pylians3_amd import Pk_library as PKL
    from pylians3_amd.batch import Pk_batch
    nreal, dims, box = 5, 32, 200.0
    rng = np.random.default_rng(11)
    deltas = rng.standard_normal((nreal, 1, dims, dims, dims)).astype(np.float32)
    res = Pk_batch(deltas, box, axis=axis, MAS=['CIC'])
    for r in range(nreal):
        single = PKL.Pk(deltas[r, 0], box, axis=axis, MAS='CIC', verbose=False)
        np.testing.assert_array_equal(res[r].Pk, single.Pk)
        np.testing.assert_array_equal(res[r].Pk2D, single.Pk2D)
        np.testing.assert_array_equal(res[r].Pk1D, single.Pk1D)
        np.testing.assert_array_equal(res[r].Nmodes3D, single.Nmodes3D)
        ref = oracle.pk(deltas[r, 0], box, axis=axis, MAS='CIC')
        assert_bins(res[r].Pk[:, 0], ref.Pk[:, 0], ref.Pk[:, 0])


def test_xpk_batch_with_prologue(oracle):
    from pylians3_amd import Pk_library as PKL
    from pylians3_amd.batch import Pk_batch
    nreal, dims, box = 4, 32, 500.0
    rng = np.random.default_rng(12)
    grids = (rng.random((nreal, 2, dims, dims, dims)) * 3 + 0.5).astype(np.float32)
    means = grids.mean(axis=(2, 3, 4)).astype(np.float32)
    res = Pk_batch(grids, box, axis=2, MAS=['CIC', 'TSC'], scales=means,
                   shifts=np.ones((nreal, 2), np.float32))
    for r in range(nreal):
        deltas = [grids[r, f] / means[r, f] - np.float32(1.0) for f in range(2)]
        single = PKL.XPk(deltas, box, axis=2, MAS=['CIC', 'TSC'])
        np.testing.assert_array_equal(res[r].Pk, single.Pk)
        np.testing.assert_array_equal(res[r].XPk, single.XPk)
        np.testing.assert_array_equal(res[r].PkX2D, single.PkX2D)
        ref = oracle.xpk(deltas, box, axis=2, MAS=['CIC', 'TSC'])
        assert_bins(res[r].Pk[:, 0, 1], ref.Pk[:, 0, 1], ref.Pk[:, 0, 1])
        fl = np.sqrt(np.abs(ref.Pk[:, 0, 0] * ref.Pk[:, 0, 1]))
        assert_bins(res[r].XPk[:, 0, 0], ref.XPk[:, 0, 0], fl)


def test_batch_device_tensors(oracle):
    import torch
    from pylians3_amd.batch import MA_batch, Pk_batch
    nreal, dims, box = 6, 32, 100.0
    offs, pos, W = _realisations(nreal, dims, box, 13)
    tpos = torch.from_numpy(pos).cuda()
    tW = torch.from_numpy(W).cuda()
    grids = torch.zeros((nreal, dims, dims, dims), dtype=torch.float32, device='cuda')
    MA_batch(tpos, offs, grids, box, 'CIC', tW, accumulate=False)
    host = np.zeros((nreal, dims, dims, dims), np.float32)
    MA_batch(pos, offs, host, box, 'CIC', W)
    np.testing.assert_array_equal(grids.cpu().numpy(), host)
    res = Pk_batch(grids.view(nreal, 1, dims, dims, dims).clone(), box, axis=1, MAS=['CIC'],
                   destroy=True)
    for r in (0, 3):
        if host[r].sum() == 0:
            continue
        ref = oracle.pk(host[r], box, axis=1, MAS='CIC')
        assert_bins(res[r].Pk[:, 0], ref.Pk[:, 0], ref.Pk[:, 0])
