"""GPU parity of the redshift-space shift (redshift_space_library.pyx:29-46) and of
its fusion into the mass assignment.

pos_redshift_space on the GPU is bitwise equal to the oracle's restatement
(float32 shift, "while < 0: += L", "if > L: fmod"), including particles that
cross the box several times and the x == L edge.  MA_rsd (shift fused into the
MA load) is bitwise equal to pos_redshift_space followed by MA, for every scheme,
accumulating and with the density-contrast output; against the oracle's shift +
serial float32 MA per cell within 1e-5 (the tolerance of the MA tests).
"""
import numpy as np
import pytest

from conftest import assert_cells, contrib_counts

pytestmark = pytest.mark.gpu


def _pv(n, box, seed, vscale=3000.0):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    vel = rng.normal(0.0, vscale, (n, 3)).astype(np.float32)
    # edge cases: several box crossings, landing exactly on 0 and L
    pos[:4] = np.float32(box) * np.float32(0.5)
    vel[:4, :] = np.array([[-1e6] * 3, [5e5] * 3, [0.0] * 3, [-50000.0] * 3], np.float32)
    pos[4] = np.float32(box)
    vel[4] = 0.0
    return pos, vel


@pytest.mark.parametrize('axis', [0, 1, 2])
def test_pos_redshift_space_vs_oracle(oracle, axis):
    import torch
    from pylians3_amd import redshift_space_library as RSL
    box, H, z = 500.0, 100.0, 0.5
    pos, vel = _pv(100000, box, axis)
    ref = oracle.pos_redshift_space(pos.copy(), vel, box, H, z, axis)
    got = pos.copy()
    RSL.pos_redshift_space(got, vel, box, H, z, axis)
    np.testing.assert_array_equal(got, ref)
    tp = torch.from_numpy(pos).cuda()
    RSL.pos_redshift_space(tp, torch.from_numpy(vel).cuda(), box, H, z, axis)
    np.testing.assert_array_equal(tp.cpu().numpy(), ref)


@pytest.mark.parametrize('mas', ['NGP', 'CIC', 'TSC', 'PCS'])
def test_ma_rsd_fused_equals_shift_then_ma(oracle, mas):
    import torch
    from pylians3_amd import MAS_library as MASL
    from pylians3_amd import redshift_space_library as RSL
    print('case ma_rsd', mas, flush=True)
    box, H, z, axis, d = 300.0, 70.0, 1.0, 0, 48
    pos, vel = _pv(200000, box, 11)
    W = np.random.default_rng(12).random(pos.shape[0]).astype(np.float32)
    tp, tv, tw = (torch.from_numpy(a).cuda() for a in (pos, vel, W))
    # fused, accumulating onto an existing grid
    init = np.random.default_rng(13).random((d,) * 3).astype(np.float32)
    fused = torch.from_numpy(init.copy()).cuda()
    MASL.MA_rsd(tp, tv, fused, box, H, z, axis, MAS=mas, W=tw)
    # two steps
    shifted = tp.clone()
    RSL.pos_redshift_space(shifted, tv, box, H, z, axis)
    two = torch.from_numpy(init.copy()).cuda()
    MASL.MA(shifted, two, box, mas, tw)
    np.testing.assert_array_equal(fused.cpu().numpy(), two.cpu().numpy())
    # host arrays take the same path
    host = init.copy()
    MASL.MA_rsd(pos, vel, host, box, H, z, axis, MAS=mas, W=W)
    np.testing.assert_array_equal(host, two.cpu().numpy())
    # against the oracle: shift, then the serial float32 MA
    sp = oracle.pos_redshift_space(pos.copy(), vel, box, H, z, axis)
    ref = init.copy()
    oracle.ma_f32(sp, ref, box, mas, W)
    assert_cells(host, ref, what='MA_rsd %s' % mas, nadds=contrib_counts(sp, d, box, mas))


def test_ma_rsd_delta_and_atomic(oracle):
    import ctypes
    import torch
    from pylians3_amd import MAS_library as MASL
    from pylians3_amd import _native as N
    box, H, z, axis, d = 300.0, 70.0, 0.0, 2, 40
    pos, vel = _pv(150000, box, 21)
    tp, tv = torch.from_numpy(pos).cuda(), torch.from_numpy(vel).cuda()
    mean = np.float32(pos.shape[0] / d ** 3)
    delta = torch.empty((d,) * 3, dtype=torch.float32, device='cuda')
    MASL.MA_rsd(tp, tv, delta, box, H, z, axis, MAS='TSC', overwrite=True, scale=float(mean),
                shift=1.0)
    grid = torch.zeros((d,) * 3, dtype=torch.float32, device='cuda')
    MASL.MA_rsd(tp, tv, grid, box, H, z, axis, MAS='TSC')
    np.testing.assert_array_equal(delta.cpu().numpy(), grid.cpu().numpy() / mean - np.float32(1))
    # the plain atomic form with the fused shift agrees to float32 summation error
    ctx = N.default_context()
    at = torch.zeros((d,) * 3, dtype=torch.float32, device='cuda')
    with N.torch_stream(ctx, True):
        N.check(N.lib().pyl_ma_rsd(ctx.handle, N.device_ptr(tp), N.device_ptr(tv), pos.shape[0],
                                   None, N.device_ptr(at), d, box, N.TSC, H, z, axis, 1.0, 0.0,
                                   N.PTR_DEVICE | N.MA_ATOMIC))
    torch.cuda.synchronize()
    sp = oracle.pos_redshift_space(pos.copy(), vel, box, H, z, axis)
    assert_cells(at.cpu().numpy(), grid.cpu().numpy(), what='atomic rsd',
                 nadds=contrib_counts(sp, d, box, 'TSC'))


def test_ma_rsd_rejects_bad_arguments():
    import ctypes
    from pylians3_amd import _native as N
    ctx = N.default_context()
    pos = np.zeros((4, 3), np.float32)
    g = np.zeros((8,) * 3, np.float32)
    st = N.lib().pyl_ma_rsd(ctx.handle, pos.ctypes.data, None, 4, None, g.ctypes.data, 8, 10.0,
                            N.CIC, 100.0, 0.0, 0, 1.0, 0.0, 0)
    assert st == N.PYL_EINVAL
    st = N.lib().pyl_ma_rsd(ctx.handle, pos.ctypes.data, pos.ctypes.data, 4, None, g.ctypes.data,
                            8, 10.0, N.CIC, 100.0, 0.0, 3, 1.0, 0.0, 0)
    assert st == N.PYL_EINVAL
