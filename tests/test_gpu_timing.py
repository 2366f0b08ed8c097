"""Stage timing of the C ABI (pyl_ctx_enable_timing / stage_times / stage_totals).

Totals accumulate over calls without a host synchronisation between them and are
read back after the stream; counts are one per stage per call.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_stage_totals_count_every_call():
    import torch
    from pylians3_amd import _native as N
    print('case stage totals', flush=True)
    ctx = N.Context(0)
    try:
        d, box, n = 64, 100.0, 200000
        pos = torch.from_numpy((np.random.default_rng(0).random((n, 3)) * box).astype(np.float32)).cuda()
        grid = torch.empty((d,) * 3, dtype=torch.float32, device='cuda')
        out, arrs, _ = N.alloc_pk_out(d, 1)
        ptrs = (ctypes.c_void_p * 1)(grid.data_ptr())
        mas = (ctypes.c_int * 1)(N.CIC)
        ctx.enable_timing(True)
        calls = 70                          # more than the event ring holds
        for _ in range(calls):
            N.check(N.lib().pyl_ma(ctx.handle, N.device_ptr(pos), n, 3, None, N.device_ptr(grid), d,
                                   box, N.CIC, N.PTR_DEVICE | N.MA_OVERWRITE | N.NO_SYNC))
            N.check(N.lib().pyl_pk(ctx.handle, ptrs, 1, mas, d, box, 0, 1.0, 0.0, N.PTR_DEVICE,
                                   ctypes.byref(out)))
        last = ctx.stage_times()
        tot = ctx.stage_totals()
        for st in ('ma_count', 'ma_scatter', 'ma_deposit', 'ma_final', 'pk_fft', 'pk_bin'):
            ms, cnt = tot[st]
            assert cnt == calls, (st, cnt)
            assert ms > 0.0 and np.isfinite(ms), (st, ms)
        assert tot['ma_atomic'][1] == 0
        assert last['pk_fft'] > 0.0 and last['ma_atomic'] == 0.0
        ctx.enable_timing(True)             # clears the totals
        assert all(c == 0 for _, c in ctx.stage_totals().values())
        # a stage mask records only the named stages
        ctx.enable_timing(True, stages=['ma_deposit'])
        for _ in range(3):
            N.check(N.lib().pyl_ma(ctx.handle, N.device_ptr(pos), n, 3, None, N.device_ptr(grid), d,
                                   box, N.CIC, N.PTR_DEVICE | N.MA_OVERWRITE | N.NO_SYNC))
        tot = ctx.stage_totals()
        assert tot['ma_deposit'][1] == 3 and tot['ma_deposit'][0] > 0.0
        assert all(c == 0 for st, (_, c) in tot.items() if st != 'ma_deposit')
        del arrs
    finally:
        ctx.close()
