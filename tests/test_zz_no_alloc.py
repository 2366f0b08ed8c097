"""pyl_ma_noalloc (fixed scratch, no allocation) against pyl_ma through the C ABI.

Both run the tiled algorithm, whose grid is bitwise reproducible, so the two
outputs must be identical.  Two calls in a row, the second with a smaller input
(fewer particles, a smaller grid), check that state left in the shared fixed
array by one call does not leak into the next.  A request beyond the array's
size is refused before anything runs.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, ctx, pos, W, grid, dims, box, mas, flags):
    from pylians3_amd import _native as N
    with N.torch_stream(ctx, True):
        st = fn(ctx.handle, N.device_ptr(pos), pos.shape[0], pos.shape[1],
                None if W is None else N.device_ptr(W), N.device_ptr(grid), dims,
                ctypes.c_float(box), mas, flags | N.PTR_DEVICE)
    return st


@pytest.mark.parametrize('mas', [1, 2, 3, 4])
@pytest.mark.parametrize('weighted', [False, True])
def test_noalloc_matches_pyl_ma(mas, weighted):
    import torch
    from pylians3_amd import _native as N
    L = N.lib()
    ctx = N.Context(0)
    try:
        cases = [(96, 300.0, 1_500_000, 3), (40, 120.0, 200_000, 3), (64, 50.0, 300_000, 2)]
        for i, (dims, box, n, axes) in enumerate(cases):
            print('case noalloc mas=%d w=%s call %d dims=%d n=%d axes=%d'
                  % (mas, weighted, i, dims, n, axes), flush=True)
            g = torch.Generator(device='cuda').manual_seed(100 * mas + i)
            pos = torch.rand((n, axes), generator=g, device='cuda') * box
            W = (torch.rand(n, generator=g, device='cuda') * 3 + 0.1) if weighted else None
            shape = (dims,) * axes
            init = torch.rand(shape, generator=g, device='cuda')
            a = init.clone()
            b = init.clone()
            flags = N.MA_CYTHON_2D | N.MA_RENORM_2D if axes == 2 else 0
            N.check(_call(L.pyl_ma, ctx, pos, W, a, dims, box, mas, flags))
            N.check(_call(L.pyl_ma_noalloc, ctx, pos, W, b, dims, box, mas, flags))
            torch.cuda.synchronize()
            assert torch.equal(a, b), 'noalloc differs from pyl_ma (call %d)' % i
            # overwrite mode, into a grid full of garbage
            b.fill_(7.0)
            a.zero_()
            N.check(_call(L.pyl_ma, ctx, pos, W, a, dims, box, mas, flags | N.MA_OVERWRITE))
            N.check(_call(L.pyl_ma_noalloc, ctx, pos, W, b, dims, box, mas, flags | N.MA_OVERWRITE))
            torch.cuda.synchronize()
            assert torch.equal(a, b), 'noalloc overwrite differs from pyl_ma (call %d)' % i
    finally:
        ctx.close()


def test_noalloc_rejects_oversize_and_host_pointers():
    import torch
    from pylians3_amd import _native as N
    L = N.lib()
    ctx = N.Context(0)
    try:
        n = 12_000_000                    # 144 MB of records alone
        pos = torch.zeros((n, 3), device='cuda')
        grid = torch.full((32, 32, 32), 3.0, device='cuda')
        st = _call(L.pyl_ma_noalloc, ctx, pos, None, grid, 32, 10.0, 2, 0)
        assert st == N.PYL_EBUDGET, st
        assert 'fixed scratch' in N.last_error()
        torch.cuda.synchronize()
        assert bool((grid == 3.0).all())  # nothing ran
        h = np.zeros((10, 3), np.float32)
        hg = np.zeros((8, 8, 8), np.float32)
        st = L.pyl_ma_noalloc(ctx.handle, h.ctypes.data, 10, 3, None, hg.ctypes.data, 8,
                              ctypes.c_float(1.0), 2, 0)
        assert st == N.PYL_EINVAL
    finally:
        ctx.close()
