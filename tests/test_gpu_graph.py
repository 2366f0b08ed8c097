"""Batched MA + Pk/XPk captured as one HIP graph (batch.BatchGraph) and the tuner.

* The graph's grids and spectra equal the uncaptured batched calls
  (pyl_ma_batch + pyl_pk_batch) bitwise: the sort chunks differ, but every tile
  sum is an exact integer and the binning order is the same.
* Replays follow new inputs written in place: other positions, other counts per
  realisation (empty ones included), other weights; each replay is checked, with
  eager work and synchronisations between the replays.
* A realisation of a replay matches the oracle (MAS_library.pyx:123-216 CIC/CICW,
  Pk_library.pyx:529-793 XPk).
* The tuner tries its candidates, caches the best and reads it back.
"""
import ctypes

import numpy as np
import pytest

from conftest import assert_bins, assert_cells, contrib_counts

pytestmark = pytest.mark.gpu


def _inputs(nreal, cap_t, box, seed, dims=32):
    import torch
    g = torch.Generator(device='cuda').manual_seed(seed)
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, cap_t // nreal, nreal)
    counts[2] = 0
    offs_t = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    tpos = torch.rand((cap_t, 3), generator=g, device='cuda') * box
    tw = torch.exp(torch.randn(cap_t, generator=g, device='cuda'))
    nm = dims ** 3 // 2
    offs_m = np.arange(nreal + 1, dtype=np.int64) * nm
    mpos = torch.rand((nreal * nm, 3), generator=g, device='cuda') * box
    return tpos, tw, offs_t, mpos, offs_m


def _uncaptured(tpos, tw, offs_t, mpos, offs_m, nreal, dims, box, scales):
    import torch
    from pylians3_amd import batch
    gt = torch.empty((nreal, dims, dims, dims), device='cuda')
    gm = torch.empty_like(gt)
    batch.MA_batch(tpos, offs_t, gt, box, 'CIC', W=tw, accumulate=False)
    batch.MA_batch(mpos, offs_m, gm, box, 'TSC', accumulate=False)
    grids = (gt.clone(), gm.clone())
    res = batch.Pk_batch([gt, gm], box, axis=1, MAS=('CIC', 'TSC'), scales=scales,
                         shifts=np.ones_like(scales), destroy=True)
    return grids, res


def test_graph_equals_uncaptured_and_follows_inputs(oracle):
    import torch
    from pylians3_amd.batch import BatchGraph
    nreal, dims, box = 6, 32, 150.0
    cap_t = 120_000
    tpos, tw, offs_t, mpos, offs_m = _inputs(nreal, cap_t, box, 1, dims)
    d_offs_t = torch.from_numpy(offs_t).cuda()
    d_offs_m = torch.from_numpy(offs_m).cuda()
    fields = [dict(pos=tpos, offs=d_offs_t, W=tw, mas='CIC'),
              dict(pos=mpos, offs=d_offs_m, W=None, mas='TSC')]
    g = BatchGraph(fields, nreal, dims, box, axis=1, tuning=dict(chunks_per_real=8))
    try:
        for rep in range(3):
            print('case graph replay', rep, flush=True)
            if rep > 0:
                # new particles and counts in place, then eager work between replays
                t2, w2, o2, m2, _ = _inputs(nreal, cap_t, box, 10 + rep, dims)
                tpos.copy_(t2)
                tw.copy_(w2)
                mpos.copy_(m2)
                offs_t = o2
                d_offs_t.copy_(torch.from_numpy(o2).cuda())
                torch.cuda.synchronize()
            scales = g.means().cpu().numpy()
            res = g.run(scales=scales)
            (ut, um), ures = _uncaptured(tpos, tw, offs_t, mpos, offs_m, nreal, dims, box, scales)
            for r in range(nreal):
                np.testing.assert_array_equal(res[r].Pk, ures[r].Pk, err_msg='replay %d r=%d' % (rep, r))
                np.testing.assert_array_equal(res[r].XPk, ures[r].XPk)
                np.testing.assert_array_equal(res[r].Pk2D, ures[r].Pk2D)
                np.testing.assert_array_equal(res[r].PkX1D, ures[r].PkX1D)
                np.testing.assert_array_equal(res[r].Nmodes3D, ures[r].Nmodes3D)
            # one realisation of this replay against the oracle
            r = 3
            t0, t1 = offs_t[r], offs_t[r + 1]
            tp = tpos[t0:t1].cpu().numpy()
            tww = tw[t0:t1].cpu().numpy()
            mp = mpos[offs_m[r]:offs_m[r + 1]].cpu().numpy()
            ref_t = np.zeros((dims,) * 3, np.float32)
            oracle.ma_f32(tp, ref_t, box, 'CIC', tww)
            assert_cells(ut[r].cpu().numpy(), ref_t, nadds=contrib_counts(tp, dims, box, 'CIC'))
            ref_m = np.zeros((dims,) * 3, np.float32)
            oracle.ma_f32(mp, ref_m, box, 'TSC')
            assert_cells(um[r].cpu().numpy(), ref_m, nadds=contrib_counts(mp, dims, box, 'TSC'))
            d_t = ut[r].cpu().numpy() / np.float32(scales[r, 0]) - np.float32(1.0)
            d_m = um[r].cpu().numpy() / np.float32(scales[r, 1]) - np.float32(1.0)
            ref = oracle.xpk([d_t, d_m], box, axis=1, MAS=['CIC', 'TSC'])
            for f in range(2):
                assert_bins(res[r].Pk[:, 0, f], ref.Pk[:, 0, f], ref.Pk[:, 0, f], what='graph auto %d' % f)
            fl = np.sqrt(np.abs(ref.Pk[:, 0, 0] * ref.Pk[:, 0, 1]))
            assert_bins(res[r].XPk[:, 0, 0], ref.XPk[:, 0, 0], fl, what='graph cross')
    finally:
        g.close()


def test_graph_rejects_growth_during_capture():
    """A call that would create a plan while capturing fails cleanly, and the
    context works again after the capture is abandoned."""
    import torch
    from pylians3_amd import _native as N
    L = N.lib()
    ctx = N.Context(0)
    try:
        delta = torch.rand((64, 64, 64), device='cuda')
        N.check(L.pyl_graph_begin(ctx.handle))
        out, arrs, _ = N.alloc_pk_out(64, 1)
        st = L.pyl_pk(ctx.handle, (ctypes.c_void_p * 1)(delta.data_ptr()), 1, (ctypes.c_int * 1)(2), 64,
                      ctypes.c_double(100.0), 0, ctypes.c_float(1.0), ctypes.c_float(0.0),
                      N.PTR_DEVICE, ctypes.byref(out))
        assert st != N.PYL_OK and 'captured' in N.last_error()
        N.check(L.pyl_graph_abort(ctx.handle))
        N.check(L.pyl_pk(ctx.handle, (ctypes.c_void_p * 1)(delta.data_ptr()), 1, (ctypes.c_int * 1)(2), 64,
                         ctypes.c_double(100.0), 0, ctypes.c_float(1.0), ctypes.c_float(0.0),
                         N.PTR_DEVICE, ctypes.byref(out)))
        assert np.all(np.isfinite(arrs['Pk3D']))
    finally:
        ctx.close()


def test_tuner_caches_its_choice(tmp_path, monkeypatch):
    import torch
    from pylians3_amd import tune
    from pylians3_amd.batch import BatchGraph
    monkeypatch.setenv('PYLIANS3_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    nreal, dims, box = 4, 32, 100.0
    tpos, tw, offs_t, mpos, offs_m = _inputs(nreal, 40_000, box, 3, dims)
    fields = [dict(pos=tpos, offs=torch.from_numpy(offs_t).cuda(), W=tw, mas='CIC')]
    cand = {'chunks_per_real': (2, 8), 'bin_rows': (0, 8), 'deposit_order': (-1,)}
    best = tune.tune_batch(fields, nreal, dims, box, 0, candidates=cand, reps=2)
    assert not best['cached'] and len(best['timings']) == 4
    again = tune.tune_batch(fields, nreal, dims, box, 0, candidates=cand, reps=2)
    assert again['cached'] and again['chunks_per_real'] == best['chunks_per_real']
    g = BatchGraph(fields, nreal, dims, box, 0, tuning={k: again[k] for k in cand})
    try:
        res = g.run()
        assert len(res) == nreal and np.all(np.isfinite(res[0].Pk))
    finally:
        g.close()
    # the tuned parameters change no bit: tile and bin sums are exact integers
    outs = []
    for tuning in (dict(chunks_per_real=2, bin_rows=0, deposit_order=-1),
                   dict(chunks_per_real=8, bin_rows=8, deposit_order=0)):
        g = BatchGraph(fields, nreal, dims, box, 0, tuning=tuning)
        try:
            outs.append(g.run())
        finally:
            g.close()
    for r in range(nreal):
        for a, b in ((outs[0][r], outs[1][r]), (outs[0][r], res[r])):
            np.testing.assert_array_equal(a.Pk, b.Pk)
            np.testing.assert_array_equal(a.Pk2D, b.Pk2D)
            np.testing.assert_array_equal(a.Pk1D, b.Pk1D)
