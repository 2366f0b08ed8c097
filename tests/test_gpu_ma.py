"""GPU parity of the mass-assignment path (libpylians_hip.so) against the oracle.

Per cell: |g - g_ref| <= 1e-5 * max(|g_ref|, mean|g_ref|) against the serial
float32 restatement of MAS_library.pyx:57-545 (bitwise equal to the reference's
MAS_c.c, see tests/test_oracle_ma.py), for every scheme, with and without
weights, 2D and 3D, both GPU algorithms.  The tiled algorithm is also checked
for bitwise run-to-run reproducibility and against the float64 oracle.
"""
import numpy as np
import pytest

from conftest import assert_cells, contrib_counts

pytestmark = pytest.mark.gpu

SCHEMES = ['NGP', 'CIC', 'TSC', 'PCS']


def _pos(n, coord, box, seed):
    return (np.random.default_rng(seed).random((n, coord)) * box).astype(np.float32)


def _gpu_ma(pos, dims, box, mas, W=None, atomic=False, init=None, renorm=True):
    from pylians3_amd import MAS_library as MASL
    coord = pos.shape[1]
    g = np.zeros((dims,) * coord, np.float32) if init is None else init.copy()
    if atomic:
        MASL.MA_atomic(pos, g, box, mas, W)
    else:
        MASL.MA(pos, g, box, mas, W, renormalize_2D=renorm)
    return g


@pytest.mark.parametrize('mas', SCHEMES)
@pytest.mark.parametrize('weighted', [False, True])
@pytest.mark.parametrize('atomic', [False, True])
def test_ma_3d_vs_oracle(oracle, mas, weighted, atomic):
    print('case ma3d', mas, weighted, atomic, flush=True)
    box, dims, n = 100.0, 40, 60000
    pos = _pos(n, 3, box, 1)
    W = (np.random.default_rng(2).lognormal(0, 1, n)).astype(np.float32) if weighted else None
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, mas, W)
    g = _gpu_ma(pos, dims, box, mas, W, atomic)
    assert_cells(g, ref, what='%s w=%s atomic=%s' % (mas, weighted, atomic),
                 nadds=contrib_counts(pos, dims, box, mas))


@pytest.mark.parametrize('mas', SCHEMES)
@pytest.mark.parametrize('weighted', [False, True])
@pytest.mark.parametrize('renorm', [True, False])
def test_ma_2d_vs_oracle(oracle, mas, weighted, renorm):
    print('case ma2d', mas, weighted, renorm, flush=True)
    box, dims, n = 50.0, 72, 40000
    pos = _pos(n, 2, box, 3)
    W = (np.random.default_rng(4).random(n) * 2).astype(np.float32) if weighted else None
    init = np.random.default_rng(5).random((dims, dims)).astype(np.float32)
    ref = init.copy()
    oracle.ma_f32(pos, ref, box, mas, W, cython2d=True, renormalize_2D=renorm)
    g = _gpu_ma(pos, dims, box, mas, W, init=init, renorm=renorm)
    assert_cells(g, ref, what='2D %s' % mas, nadds=contrib_counts(pos, dims, box, mas))


@pytest.mark.parametrize('dims', [2, 3, 5, 16, 17, 20, 33, 97, 100])
@pytest.mark.parametrize('mas', SCHEMES)
def test_ma_grid_sizes_and_tile_edges(oracle, dims, mas):
    print('case sizes', dims, mas, flush=True)
    box = 10.0
    pos = _pos(20000, 3, box, dims)
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, mas)
    assert_cells(_gpu_ma(pos, dims, box, mas), ref, what='dims=%d %s' % (dims, mas),
                 nadds=contrib_counts(pos, dims, box, mas))


@pytest.mark.parametrize('mas', SCHEMES)
def test_ma_edge_positions(oracle, mas):
    print('case edges', mas, flush=True)
    box, dims = 64.0, 32
    # inv = dims/box = 0.5: x = 2k+1 puts dist exactly on k + 0.5 (the TSC/NGP
    # branch points), its float neighbours just either side
    half = np.float32(7.0)
    vals = np.array([0.0, box, np.nextafter(np.float32(box), np.float32(0)), 1.0, 2.0, 2.5,
                     31.999, 32.0, 63.0, half, np.nextafter(half, np.float32(0)),
                     np.nextafter(half, np.float32(100)), np.float32(1e-8)], np.float32)
    g = np.stack(np.meshgrid(vals, vals, vals, indexing='ij'), -1).reshape(-1, 3)
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(g, ref, box, mas)
    assert_cells(_gpu_ma(g, dims, box, mas), ref, what='edges %s' % mas)


def test_ma_empty_and_single(oracle):
    box, dims = 10.0, 16
    g = _gpu_ma(np.zeros((0, 3), np.float32), dims, box, 'CIC', init=np.ones((dims,) * 3, np.float32))
    np.testing.assert_array_equal(g, np.ones((dims,) * 3, np.float32))
    p = np.array([[3.3, 7.1, 9.99]], np.float32)
    for mas in SCHEMES:
        ref = np.zeros((dims,) * 3, np.float32)
        oracle.ma_f32(p, ref, box, mas)
        assert_cells(_gpu_ma(p, dims, box, mas), ref, what='single %s' % mas)


@pytest.mark.parametrize('mas', SCHEMES)
def test_ma_clustered_one_cell(oracle, mas):
    # every particle in one cell: one tile holds all records
    print('case clustered', mas, flush=True)
    box, dims, n = 100.0, 32, 200000
    rng = np.random.default_rng(8)
    pos = (50.0 + rng.random((n, 3)) * (box / dims)).astype(np.float32)
    W = rng.random(n).astype(np.float32)
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, mas, W)
    g64 = oracle.ma_f64(pos, dims, box, mas, W)
    g = _gpu_ma(pos, dims, box, mas, W)
    # ~1e5 additions land in the densest cells: the reference's serial float32
    # sum drifts by up to n u |g| there, the tiled sum is rounded once
    assert_cells(g, ref, what='clustered %s' % mas, nadds=contrib_counts(pos, dims, box, mas))
    err64 = np.abs(g - g64) / np.maximum(np.abs(g64), 1e-30)
    err_ref = np.abs(ref - g64) / np.maximum(np.abs(g64), 1e-30)
    # vs float64: what remains is the float32 rounding of dist = x*inv in the
    # reference arithmetic (~1e-6 of a weight), never worse than the reference's
    dense = g64 > 0.1 * g64.max()
    assert err64[dense].max() < 5e-6, err64[dense].max()
    assert err64[dense].max() <= err_ref[dense].max()


def test_ma_accumulates_and_is_reproducible(oracle):
    box, dims, n = 100.0, 48, 300000
    rng = np.random.default_rng(10)
    pos = (rng.normal(50, 8, (n, 3)) % box).astype(np.float32)
    init = rng.random((dims,) * 3).astype(np.float32)
    a = _gpu_ma(pos, dims, box, 'TSC', init=init)
    b = _gpu_ma(pos, dims, box, 'TSC', init=init)
    np.testing.assert_array_equal(a, b)
    ref = init.copy()
    oracle.ma_f32(pos, ref, box, 'TSC')
    assert_cells(a, ref, nadds=contrib_counts(pos, dims, box, 'TSC'))


def test_ma_device_tensors(oracle):
    import torch
    box, dims, n = 100.0, 32, 50000
    pos = _pos(n, 3, box, 12)
    W = np.random.default_rng(13).random(n).astype(np.float32)
    from pylians3_amd import MAS_library as MASL
    tp = torch.from_numpy(pos).cuda()
    tw = torch.from_numpy(W).cuda()
    tg = torch.zeros((dims,) * 3, dtype=torch.float32, device='cuda')
    MASL.MA(tp, tg, box, 'PCS', tw)
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, 'PCS', W)
    assert_cells(tg.cpu().numpy(), ref)


@pytest.mark.parametrize('mas', SCHEMES)
def test_mas_c_abi_names(oracle, mas):
    # NGP/CIC/TSC/PCS exported with the MAS_c.h:3-10 signatures (native 2D)
    import ctypes
    from pylians3_amd import _native as N
    box = 20.0
    for coord, dims in ((3, 24), (2, 40)):
        pos = _pos(30000, coord, box, 14 + coord)
        W = np.random.default_rng(15).random(30000).astype(np.float32)
        g = np.zeros((dims,) * coord, np.float32)
        getattr(N.lib(), mas)(pos.ctypes.data, g.ctypes.data, W.ctypes.data, ctypes.c_long(30000),
                              dims, coord, ctypes.c_float(box), 4)
        ref = np.zeros((dims,) * coord, np.float32)
        oracle.ma_f32(pos, ref, box, mas, W, cython2d=False)
        assert_cells(g, ref, what='C ABI %s %dD' % (mas, coord))


@pytest.mark.slow
@pytest.mark.parametrize('mas', SCHEMES)
def test_ma_large_mass_conservation(mas):
    import torch
    from pylians3_amd import MAS_library as MASL
    dims, n = 256, 1 << 24
    g = torch.Generator(device='cuda').manual_seed(0)
    pos = torch.rand((n, 3), generator=g, device='cuda', dtype=torch.float32) * 1000.0
    grid = torch.zeros((dims,) * 3, dtype=torch.float32, device='cuda')
    MASL.MA(pos, grid, 1000.0, mas)
    tot = grid.double().sum().item()
    assert abs(tot / n - 1.0) < 1e-5
    grid2 = torch.zeros_like(grid)
    MASL.MA(pos, grid2, 1000.0, mas)
    assert torch.equal(grid, grid2)


@pytest.mark.parametrize('mas', ['CIC', 'TSC'])
def test_ma_ordered_tiles_vs_oracle(oracle, mas):
    # 256^3 grid (4096 tiles) with 4M clustered particles, 1024 per tile on
    # average: the deposit takes its tiles largest first; every cell matches the
    # oracle and the result is bitwise reproducible
    print('case ordered tiles', mas, flush=True)
    box, dims, n = 1000.0, 256, 1 << 22
    rng = np.random.default_rng(77)
    centres = rng.random((400, 3)) * box
    pos = centres[rng.integers(0, 400, n)] + rng.normal(0.0, 8.0, (n, 3))
    pos = np.mod(pos, box).astype(np.float32)
    pos[pos >= box] = 0.0
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, mas)
    g = _gpu_ma(pos, dims, box, mas)
    assert_cells(g, ref, what='ordered tiles %s' % mas, nadds=contrib_counts(pos, dims, box, mas))
    assert np.array_equal(g, _gpu_ma(pos, dims, box, mas))


@pytest.mark.parametrize('mas', SCHEMES)
def test_ma_mass_weighted_sparse_vs_oracle(oracle, mas):
    # halo-catalogue shape: 200k positions on 256^3 (one per 84 cells) weighted by
    # masses spanning five decades (1e10-1e15, log-uniform), and signed weights
    # (velocities): with per-particle weights each contribution is held in 64-bit
    # fixed point scaled by n * max|W|, so even cells far below max|W| match
    print('case mass weighted sparse', mas, flush=True)
    box, dims, n = 1000.0, 256, 200_000
    rng = np.random.default_rng(41)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    pos[pos >= box] = 0.0
    nadds = contrib_counts(pos, dims, box, mas)
    for w in ((10.0 ** rng.uniform(10.0, 15.0, n)).astype(np.float32),
              rng.normal(0.0, 300.0, n).astype(np.float32)):
        ref = np.zeros((dims,) * 3, np.float32)
        oracle.ma_f32(pos, ref, box, mas, w)
        g = _gpu_ma(pos, dims, box, mas, w)
        assert_cells(g, ref, what='mass weighted %s' % mas, nadds=nadds)
        assert np.array_equal(g, _gpu_ma(pos, dims, box, mas, w))


def _forced_ma(ctx, force, pos, n, w, g, dims, box, masid):
    """pyl_ma on ctx with the sort form forced; returns the form taken."""
    import ctypes
    from pylians3_amd import _native as N
    L = N.lib()
    N.check(L.pyl_ctx_set_tuning(ctx.handle, N.TUNE_SORT_WINDOW, force))
    N.check(L.pyl_ma(ctx.handle, pos.ctypes.data, n, 3, None if w is None else w.ctypes.data,
                     g.ctypes.data, dims, ctypes.c_float(box), masid, 0))
    got = ctypes.c_int()
    N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
    return got.value


@pytest.mark.parametrize('mas', ['CIC', 'PCS'])
@pytest.mark.parametrize('dims', [512, 480])
def test_ma_wide_window_any_order(oracle, mas, dims):
    # 3D grids of 16384 < tiles <= 32768 (512^3: 32^3 tiles; 480^3: 30^3 tiles)
    # can be sorted in one pass with the 8 x 8 tile-column box window or the wide
    # window (every tile in LDS), or with the two-level sort (buckets of 8 x 8 x 4
    # tiles, then tiles).  Each form is forced in turn on a context of its own and
    # the form taken is read back; every cell matches the oracle, with and without
    # weights, in random order and sorted by cell, and all forms give the same bits.
    # Then the adaptive choice: random order moves the next call to the two-level
    # sort, tile-ordered input back to one pass, and a fresh context starts from the
    # process's latest finding for this grid and scheme.
    import ctypes
    from pylians3_amd import MAS_library as MASL, _native as N
    print('case wide window', mas, dims, flush=True)
    box, n = 1000.0, 1 << 21
    rng = np.random.default_rng(dims + 3)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    pos[pos >= box] = 0.0
    W = rng.lognormal(0.0, 1.0, n).astype(np.float32)
    nadds = contrib_counts(pos, dims, box, mas)
    cell = np.minimum((pos.astype(np.float64) * (dims / box)).astype(np.int64), dims - 1)
    srt = np.argsort((cell[:, 0] * dims + cell[:, 1]) * dims + cell[:, 2], kind='stable')
    L = N.lib()
    masid = MASL._MAS_IDS[mas]
    firsts = []
    for w in (None, W):
        ref = np.zeros((dims,) * 3, np.float32)
        oracle.ma_f32(pos, ref, box, mas, w)
        first = None
        for force, expect in ((N.SORT_ONEPASS, N.WINDOW_BOX), (N.SORT_WIDE, N.WINDOW_WIDE),
                              (N.SORT_TWOLEVEL, N.WINDOW_TWOLEVEL)):
            ctx = N.Context(0)
            try:
                for order in (None, srt):
                    p = pos if order is None else np.ascontiguousarray(pos[order])
                    ww = None if w is None else (w if order is None else np.ascontiguousarray(w[order]))
                    g = np.zeros((dims,) * 3, np.float32)
                    got = _forced_ma(ctx, force, p, n, ww, g, dims, box, masid)
                    assert got == expect, (force, got)
                    if first is None:
                        assert_cells(g, ref, what='window %d %s %d %s' % (force, mas, dims, w is not None),
                                     nadds=nadds)
                        first = g
                    else:
                        assert np.array_equal(g, first), ('window', force, order is None)
            finally:
                ctx.close()
        firsts.append(first)
    first = firsts[0]                        # unweighted
    # adaptive: a call tallies the particles outside their chunk's box; over a quarter
    # moves the next call with this (dims, MAS) to the two-level sort.  A context
    # that has not sorted such a grid yet follows the process's latest finding: the
    # calls above found random order, so a fresh context starts with two-level
    ctx = N.Context(0)
    try:
        got = ctypes.c_int()
        for k in range(3):
            g = np.zeros((dims,) * 3, np.float32)
            N.check(L.pyl_ma(ctx.handle, pos.ctypes.data, n, 3, None, g.ctypes.data, dims,
                             ctypes.c_float(box), masid, 0))
            N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
            assert got.value == N.WINDOW_TWOLEVEL, (k, got.value)
            assert np.array_equal(g, first)
        # input ordered by 16^3 tile on the same context tallies few misses (a sort
        # chunk spans a few neighbouring tile columns): back to one pass
        tl = cell // 16
        nt = -(-dims // 16)
        tsrt = np.argsort((tl[:, 0] * nt + tl[:, 1]) * nt + tl[:, 2], kind='stable')
        p = np.ascontiguousarray(pos[tsrt])
        for k in range(3):
            g = np.zeros((dims,) * 3, np.float32)
            N.check(L.pyl_ma(ctx.handle, p.ctypes.data, n, 3, None, g.ctypes.data, dims,
                             ctypes.c_float(box), masid, 0))
            assert np.array_equal(g, first)
            N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
        assert got.value == N.WINDOW_BOX
    finally:
        ctx.close()
    # the process's latest finding is now ordered input: a fresh context starts with
    # one pass, and random input moves it to two-level from its second call
    ctx = N.Context(0)
    try:
        got = ctypes.c_int()
        for k in range(3):
            g = np.zeros((dims,) * 3, np.float32)
            N.check(L.pyl_ma(ctx.handle, pos.ctypes.data, n, 3, None, g.ctypes.data, dims,
                             ctypes.c_float(box), masid, 0))
            N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
            if k == 0:
                assert got.value == N.WINDOW_BOX
            assert np.array_equal(g, first)
        assert got.value == N.WINDOW_TWOLEVEL
    finally:
        ctx.close()


@pytest.mark.parametrize('dims,mas,n,weights', [(256, 'TSC', 1 << 21, (False, True)),
                                                (400, 'PCS', 1 << 21, (False, True)),
                                                (1040, 'CIC', 1 << 22, (True,)),
                                                (1024, 'NGP', 1 << 22, (False,))])
def test_ma_two_level_sort_vs_oracle(oracle, dims, mas, n, weights):
    # The two-level sort (buckets of 8 x 8 x 4 tiles) on grids of other shapes:
    # 256^3 (16^3 tiles, every tile in the one-pass LDS window; 16 buckets), 400^3
    # (25^3 unequal tiles, partial buckets at the upper edges), 1040^3 (65^3 tiles,
    # 9 x 9 x 17 buckets, the last a single tile thick), 1024^3 (64^3 tiles, 1024
    # buckets; one pass would send random-order particles to the global overflow
    # counters).  Random order, with
    # and without weights (one of the two on the 1024-class grids): every cell
    # matches the oracle and equals the one-pass sort bitwise.
    from pylians3_amd import MAS_library as MASL, _native as N
    print('case two-level', dims, mas, n, flush=True)
    box = 700.0
    rng = np.random.default_rng(dims)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    pos[pos >= box] = 0.0
    pos[:3] = [[0.0, 0.0, 0.0], [box * 0.999999, 0.0, box * 0.5], [box * 0.5, box * 0.999999, 0.0]]
    W = rng.lognormal(0.0, 1.0, n).astype(np.float32)
    masid = MASL._MAS_IDS[mas]
    nadds = contrib_counts(pos, dims, box, mas)
    for w in [W if wt else None for wt in weights]:
        ref = np.zeros((dims,) * 3, np.float32)
        oracle.ma_f32(pos, ref, box, mas, w)
        grids = []
        for force, expect in ((N.SORT_TWOLEVEL, N.WINDOW_TWOLEVEL), (N.SORT_ONEPASS, None)):
            ctx = N.Context(0)
            try:
                g = np.zeros((dims,) * 3, np.float32)
                got = _forced_ma(ctx, force, pos, n, w, g, dims, box, masid)
                assert expect is None or got == expect, (force, got)
                grids.append(g)
            finally:
                ctx.close()
        assert_cells(grids[0], ref, what='two-level %d %s %s' % (dims, mas, w is not None), nadds=nadds)
        assert np.array_equal(grids[0], grids[1])
        del ref, grids


def test_ma_rsd_two_level_equals_one_pass():
    # the redshift-space shift is applied once, in the bucket pass
    import ctypes
    from pylians3_amd import _native as N
    print('case two-level rsd', flush=True)
    box, dims, n = 500.0, 512, 1 << 21
    rng = np.random.default_rng(77)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    vel = rng.normal(0.0, 2000.0, (n, 3)).astype(np.float32)
    W = rng.random(n).astype(np.float32)
    L = N.lib()
    out = []
    for force in (N.SORT_TWOLEVEL, N.SORT_ONEPASS):
        ctx = N.Context(0)
        try:
            N.check(L.pyl_ctx_set_tuning(ctx.handle, N.TUNE_SORT_WINDOW, force))
            g = np.zeros((dims,) * 3, np.float32)
            N.check(L.pyl_ma_rsd(ctx.handle, pos.ctypes.data, vel.ctypes.data, n, W.ctypes.data,
                                 g.ctypes.data, dims, ctypes.c_float(box), N.TSC, ctypes.c_float(100.0),
                                 ctypes.c_float(0.5), 1, ctypes.c_float(1.0), ctypes.c_float(0.0), 0))
            got = ctypes.c_int()
            N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
            assert (got.value == N.WINDOW_TWOLEVEL) == (force == N.SORT_TWOLEVEL)
            out.append(g)
        finally:
            ctx.close()
    assert np.array_equal(out[0], out[1])
    assert abs(out[0].astype(np.float64).sum() - W.astype(np.float64).sum()) <= 1e-5 * W.sum()


def _clumped_lattice(d, box, seed, frac=0.06, disp=3.0):
    """A displaced d^3 lattice in lattice order with a fraction of its particles moved
    into one sphere of 4 cells' radius (a few tiles hold many times the mean)."""
    p = _lattice(d, box, seed, disp)
    rng = np.random.default_rng(seed + 1)
    k = int(frac * p.shape[0])
    idx = np.sort(rng.choice(p.shape[0], k, replace=False))
    c = np.float32(0.37 * box)
    r = rng.normal(0.0, 4.0 * box / d / 3.0, (k, 3))
    p[idx] = ((c + r) % box).astype(np.float32) % np.float32(box)
    return p


@pytest.mark.parametrize('dims,d,mas,weighted,order', [
    (256, 256, 'TSC', False, 'lattice'),       # every tile in the window, 4096 per tile
    (256, 192, 'CIC', True, 'clumped'),        # clumped: tiles past their capacity spill
    (512, 256, 'PCS', False, 'lattice'),       # 8 x 8 tile-column box window, 512 per tile
    (512, 200, 'NGP', True, 'clumped'),
    (432, 160, 'TSC', True, 'shuffled'),       # every particle misses its chunk's box:
                                               # 4.1 M spilled, the atomic spill sort
    (432, 176, 'CIC', True, 'shuffled'),       # 5.5 M spilled (> 2^22): the windowed spill sort
    (200, 96, 'PCS', False, 'lattice'),        # unequal tiles (13 per axis)
])
def test_ma_capacity_sort_vs_oracle(oracle, dims, d, mas, weighted, order):
    # The capacity sort: one read of the particles into per-tile slots (3x the mean
    # occupancy); particles outside their chunk's window or past a full tile go to the
    # spill list, which the regular passes sort into a second segment.  Forced on a
    # context of its own: every cell matches the oracle and equals the one-pass sort
    # bitwise.
    from pylians3_amd import MAS_library as MASL, _native as N
    print('case capacity', dims, d, mas, weighted, order, flush=True)
    box = 600.0
    pos = _clumped_lattice(d, box, 40 + d) if order == 'clumped' else _lattice(d, box, 40 + d)
    if order == 'shuffled':
        pos = pos[np.random.default_rng(3).permutation(pos.shape[0])]
    pos = np.ascontiguousarray(pos)
    n = pos.shape[0]
    W = np.random.default_rng(d).lognormal(0.0, 1.0, n).astype(np.float32) if weighted else None
    masid = MASL._MAS_IDS[mas]
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, mas, W)
    grids = []
    for force, expect in ((N.SORT_CAPACITY, N.WINDOW_CAPACITY), (N.SORT_ONEPASS, None)):
        ctx = N.Context(0)
        try:
            g = np.zeros((dims,) * 3, np.float32)
            got = _forced_ma(ctx, force, pos, n, W, g, dims, box, masid)
            assert expect is None or got == expect, (force, got)
            grids.append(g)
        finally:
            ctx.close()
    assert_cells(grids[0], ref, what='capacity %d %s' % (dims, mas),
                 nadds=contrib_counts(pos, dims, box, mas))
    assert np.array_equal(grids[0], grids[1])


@pytest.mark.parametrize('dims,d,mas,weighted,clump', [
    (256, 256, 'TSC', False, False),           # 16 buckets, 4096 particles per tile
    (512, 256, 'CIC', True, True),             # 1024 buckets; clumped tiles overflow
    (400, 192, 'PCS', False, True),            # unequal tiles, partial buckets at the edges
])
def test_ma_two_level_capacity_vs_oracle(oracle, dims, d, mas, weighted, clump):
    # The two-level sort with the capacity form as its second pass: random-order
    # particles go into buckets of 8 x 8 x 4 tiles, then each bucket's records into
    # per-tile slots (only full tiles spill).  Forced on a context of its own: every
    # cell matches the oracle and equals the one-pass sort bitwise.  Then the adaptive
    # choice on such input (at least 256 particles per tile): two-level with capacity,
    # or, once a clumped input has spilled more than 1/64 of its particles, without.
    import ctypes
    from pylians3_amd import MAS_library as MASL, _native as N
    print('case two-level capacity', dims, d, mas, weighted, clump, flush=True)
    box = 600.0
    pos = _clumped_lattice(d, box, 70 + d) if clump else _lattice(d, box, 70 + d)
    pos = np.ascontiguousarray(pos[np.random.default_rng(4).permutation(pos.shape[0])])
    n = pos.shape[0]
    W = np.random.default_rng(d + 1).lognormal(0.0, 1.0, n).astype(np.float32) if weighted else None
    masid = MASL._MAS_IDS[mas]
    ref = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, mas, W)
    grids = []
    for force, expect in ((N.SORT_TWOLEVEL_CAPACITY, N.WINDOW_TWOLEVEL_CAPACITY), (N.SORT_ONEPASS, None)):
        ctx = N.Context(0)
        try:
            g = np.zeros((dims,) * 3, np.float32)
            got = _forced_ma(ctx, force, pos, n, W, g, dims, box, masid)
            assert expect is None or got == expect, (force, got)
            grids.append(g)
        finally:
            ctx.close()
    assert_cells(grids[0], ref, what='two-level capacity %d %s' % (dims, mas),
                 nadds=contrib_counts(pos, dims, box, mas))
    assert np.array_equal(grids[0], grids[1])
    if n < 256 * (-(-dims // 16)) ** 3:
        return
    ctx = N.Context(0)
    try:
        got = ctypes.c_int()
        for _ in range(3):
            g = np.zeros((dims,) * 3, np.float32)
            N.check(N.lib().pyl_ma(ctx.handle, pos.ctypes.data, n, 3, None if W is None else W.ctypes.data,
                                   g.ctypes.data, dims, ctypes.c_float(box), masid, 0))
            N.check(N.lib().pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
            assert np.array_equal(g, grids[0])
        assert got.value == (N.WINDOW_TWOLEVEL if clump else N.WINDOW_TWOLEVEL_CAPACITY)
    finally:
        ctx.close()


def test_ma_capacity_adaptive_and_spill_fallback():
    # Adaptive choice on a 256^3 grid with 256^3 lattice-ordered particles (4096 per
    # tile): the capacity sort, once the context has seen ordered input.  A call that
    # spills more than 1/64 of its particles (most particles in one clump, in cell
    # order) turns it off for that grid and scheme on the context: the next call takes
    # the one-pass sort.  All grids of the same input are bitwise equal.
    import ctypes
    from pylians3_amd import _native as N
    print('case capacity adaptive', flush=True)
    box, dims = 500.0, 256
    lat = _lattice(dims, box, 8)
    rng = np.random.default_rng(9)
    clump = ((np.float32(0.5 * box) + rng.normal(0.0, 3.0, (1 << 21, 3))) % box).astype(np.float32)
    clump = clump[np.lexsort((clump[:, 2], clump[:, 1], clump[:, 0]))]
    L = N.lib()
    ctx = N.Context(0)
    got = ctypes.c_int()

    def run(p):
        g = np.zeros((dims,) * 3, np.float32)
        N.check(L.pyl_ma(ctx.handle, p.ctypes.data, p.shape[0], 3, None, g.ctypes.data, dims,
                         ctypes.c_float(box), N.TSC, 0))
        N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
        return g, got.value

    try:
        first, _ = run(lat)
        for _ in range(2):
            g, w = run(lat)
            assert np.array_equal(g, first)
        assert w == N.WINDOW_CAPACITY
        c0, w0 = run(clump)
        assert w0 == N.WINDOW_CAPACITY
        c1, w1 = run(clump)
        assert w1 != N.WINDOW_CAPACITY
        assert np.array_equal(c0, c1)
        assert abs(c0.astype(np.float64).sum() - clump.shape[0]) <= 1e-6 * clump.shape[0]
    finally:
        ctx.close()


def test_ma_rsd_capacity_equals_one_pass():
    # the redshift-space shift is applied once, as the capacity pass reads the particles
    import ctypes
    from pylians3_amd import _native as N
    print('case capacity rsd', flush=True)
    box, dims, d = 500.0, 256, 200
    pos = _lattice(d, box, 78)
    vel = np.random.default_rng(79).normal(0.0, 300.0, pos.shape).astype(np.float32)
    W = np.random.default_rng(80).random(pos.shape[0]).astype(np.float32)
    L = N.lib()
    out = []
    for force in (N.SORT_CAPACITY, N.SORT_ONEPASS):
        ctx = N.Context(0)
        try:
            N.check(L.pyl_ctx_set_tuning(ctx.handle, N.TUNE_SORT_WINDOW, force))
            g = np.zeros((dims,) * 3, np.float32)
            N.check(L.pyl_ma_rsd(ctx.handle, pos.ctypes.data, vel.ctypes.data, pos.shape[0], W.ctypes.data,
                                 g.ctypes.data, dims, ctypes.c_float(box), N.TSC, ctypes.c_float(100.0),
                                 ctypes.c_float(0.5), 0, ctypes.c_float(1.0), ctypes.c_float(0.0), 0))
            got = ctypes.c_int()
            N.check(L.pyl_ctx_last_sort_window(ctx.handle, ctypes.byref(got)))
            assert (got.value == N.WINDOW_CAPACITY) == (force == N.SORT_CAPACITY)
            out.append(g)
        finally:
            ctx.close()
    assert np.array_equal(out[0], out[1])
    assert abs(out[0].astype(np.float64).sum() - W.astype(np.float64).sum()) <= 1e-5 * W.sum()


@pytest.mark.parametrize('mas', SCHEMES)
@pytest.mark.parametrize('dims,coord', [(32, 3), (40, 3), (128, 2)])
def test_ma_split_tiles_vs_oracle(oracle, mas, dims, coord):
    # few tiles with many records each (32^3: 8 tiles, 40^3: 27 unequal tiles, 128^2:
    # 4 tiles of 64^2): each tile's records are shared by several workgroups whose
    # integer boxes add up in a second pass; every cell matches the oracle, with and
    # without weights, and the result is bitwise reproducible
    print('case split tiles', mas, dims, coord, flush=True)
    box, n = 100.0, 1_000_000
    pos = _pos(n, coord, box, 91 + dims)
    W = np.random.default_rng(5).lognormal(0.0, 1.0, n).astype(np.float32)
    for w in (None, W):
        ref = np.zeros((dims,) * coord, np.float32)
        oracle.ma_f32(pos, ref, box, mas, w)
        g = _gpu_ma(pos, dims, box, mas, w)
        assert_cells(g, ref, what='split %s %d %s' % (mas, dims, w is not None),
                     nadds=contrib_counts(pos, dims, box, mas))
        assert np.array_equal(g, _gpu_ma(pos, dims, box, mas, w))


def test_cic_interp_vs_oracle(oracle):
    import torch
    from pylians3_amd import MAS_library as MASL
    rng = np.random.default_rng(40)
    for dims, box in ((32, 100.0), (17, 5.0)):
        g = rng.standard_normal((dims,) * 3).astype(np.float32)
        pos = (rng.random((40000, 3)) * box).astype(np.float32)
        pos[:3] = np.float32(box)
        den = np.zeros(40000, np.float32)
        MASL.CIC_interp(g, box, pos, den)
        ref = oracle.cic_interp(g, box, pos)
        np.testing.assert_allclose(den, ref, rtol=2e-6, atol=2e-6 * np.abs(g).max())
        tden = torch.zeros(40000, device='cuda')
        MASL.CIC_interp(torch.from_numpy(g).cuda(), box, torch.from_numpy(pos).cuda(), tden)
        np.testing.assert_array_equal(tden.cpu().numpy(), den)


def _lattice(d, box, seed, disp=3.0):
    """d^3 particles in lattice (C) order, displaced by disp cells rms: the coherent
    input order of a simulation snapshot."""
    rng = np.random.default_rng(seed)
    q = (np.arange(d) + 0.5) * (box / d)
    p = np.stack(np.meshgrid(q, q, q, indexing='ij'), -1).reshape(-1, 3)
    p = p + rng.normal(0.0, disp * box / d, p.shape)
    return (p % box).astype(np.float32) % np.float32(box)


@pytest.mark.parametrize('mas', SCHEMES)
def test_ma_lattice_order_vs_oracle(oracle, mas):
    # many sort chunks, coherent keys: the LDS-window counting path
    print('case lattice', mas, flush=True)
    box, d = 300.0, 80
    pos = _lattice(d, box, 21)
    ref = np.zeros((d,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, mas)
    g = _gpu_ma(pos, d, box, mas)
    assert_cells(g, ref, what='lattice %s' % mas, nadds=contrib_counts(pos, d, box, mas))


def test_ma_many_tiles_random_order(oracle):
    # more tiles than the LDS window holds, random order: most keys take the
    # global overflow counters
    import torch
    from pylians3_amd import MAS_library as MASL
    print('case overflow keys', flush=True)
    box, d, n = 500.0, 432, 150000
    pos = _pos(n, 3, box, 22)
    W = np.random.default_rng(23).random(n).astype(np.float32)
    tg = torch.zeros((d,) * 3, dtype=torch.float32, device='cuda')
    MASL.MA(torch.from_numpy(pos).cuda(), tg, box, 'TSC', torch.from_numpy(W).cuda())
    g = tg.cpu().numpy()
    del tg
    ref = np.zeros((d,) * 3, np.float32)
    oracle.ma_f32(pos, ref, box, 'TSC', W)
    assert_cells(g, ref, what='overflow keys')


@pytest.mark.parametrize('mas', ['CIC', 'TSC', 'PCS'])
def test_ma_delta_equals_ma_then_normalise(oracle, mas):
    # pyl_ma_delta == MA (overwrite) followed by "delta /= mean; delta -= 1" in float32
    import ctypes
    import torch
    from pylians3_amd import _native as N
    box, d = 200.0, 48
    pos = torch.from_numpy(_lattice(d, box, 24)).cuda()
    mean = np.float32(pos.shape[0] / d ** 3)
    ctx = N.default_context()
    a = torch.empty((d,) * 3, dtype=torch.float32, device='cuda')
    b = torch.empty_like(a)
    mid = {'CIC': N.CIC, 'TSC': N.TSC, 'PCS': N.PCS}[mas]
    with N.torch_stream(ctx, True):
        N.check(N.lib().pyl_ma(ctx.handle, N.device_ptr(pos), pos.shape[0], 3, None, N.device_ptr(a),
                               d, box, mid, N.PTR_DEVICE | N.MA_OVERWRITE))
        N.check(N.lib().pyl_ma_delta(ctx.handle, N.device_ptr(pos), pos.shape[0], 3, None,
                                     N.device_ptr(b), d, box, mid, float(mean), 1.0, N.PTR_DEVICE))
    torch.cuda.synchronize()
    ah = a.cpu().numpy()
    expect = ah / mean - np.float32(1.0)
    np.testing.assert_array_equal(b.cpu().numpy(), expect)
    # the atomic form applies the same transform
    c = torch.empty_like(a)
    with N.torch_stream(ctx, True):
        N.check(N.lib().pyl_ma_delta(ctx.handle, N.device_ptr(pos), pos.shape[0], 3, None,
                                     N.device_ptr(c), d, box, mid, float(mean), 1.0,
                                     N.PTR_DEVICE | N.MA_ATOMIC))
    torch.cuda.synchronize()
    assert_cells(c.cpu().numpy() + 1.0, expect + 1.0, what='atomic delta')


@pytest.mark.parametrize('d,sub', [(448, 1), (576, 4)])
def test_ma_box_window_lattice_equals_shuffled(d, sub):
    # 448^3 grid (28^3 tiles): every tile has an LDS counter in the wide window;
    # 576^3 (36^3 tiles, more than it holds): lattice-ordered chunks count their
    # keys in the 8 x 8 tile-column box around them and shuffled input takes the
    # global overflow counters.  The integer tile sums make lattice-ordered and
    # shuffled grids bitwise equal, and the total is the particle count
    import torch
    from pylians3_amd import _native as N
    print('case box window', d, sub, flush=True)
    box = 1000.0
    g = torch.Generator(device='cuda').manual_seed(31)
    q = (torch.arange(d, device='cuda', dtype=torch.float32) + 0.5) * (box / d)
    pos = torch.stack(torch.meshgrid(q, q, q, indexing='ij'), -1).reshape(-1, 3)
    pos = pos + torch.randn(pos.shape, generator=g, device='cuda') * (4.0 * box / d)
    pos = torch.remainder(pos, box)[::sub].contiguous()
    pos[pos >= box] = 0.0
    n = pos.shape[0]
    ctx = N.default_context()
    a = torch.empty((d,) * 3, dtype=torch.float32, device='cuda')
    b = torch.empty_like(a)
    with N.torch_stream(ctx, True):
        N.check(N.lib().pyl_ma(ctx.handle, N.device_ptr(pos), n, 3, None, N.device_ptr(a), d, box,
                               N.TSC, N.PTR_DEVICE | N.MA_OVERWRITE))
    perm = torch.randperm(n, generator=g, device='cuda')
    pos_s = pos[perm].contiguous()
    del perm, pos
    with N.torch_stream(ctx, True):
        N.check(N.lib().pyl_ma(ctx.handle, N.device_ptr(pos_s), n, 3, None, N.device_ptr(b), d, box,
                               N.TSC, N.PTR_DEVICE | N.MA_OVERWRITE))
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    tot = a.double().sum().item()
    assert abs(tot - n) <= 1e-6 * n, (tot, n)
