"""Stream ordering of calls that share one context's scratch memory.

A device MA queued on a torch side stream returns without waiting; the next
call on the same context (here a host-array MA on the context's own stream)
reuses the tile counters and records.  The context must order the new stream
after the old one, or the second call would clear counters the first call's
kernels are still reading.  Both grids are checked against the oracle.

A server request whose inputs are still being written on the submitter's
stream must not be read early either.
"""
import numpy as np
import pytest

from conftest import assert_cells, contrib_counts

pytestmark = pytest.mark.gpu


def test_side_stream_device_ma_then_host_ma(oracle):
    import torch
    from pylians3_amd import MAS_library as MASL
    box, dims = 500.0, 96
    rng = np.random.default_rng(11)
    posA = (rng.random((3_000_000, 3)) * box).astype(np.float32)
    posB = (rng.random((400_000, 3)) * box).astype(np.float32)
    refA = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(posA, refA, box, 'PCS')
    refB = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(posB, refB, box, 'CIC')
    dA = torch.from_numpy(posA).cuda()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    for it in range(3):
        print('case side-stream round', it, flush=True)
        with torch.cuda.stream(side):
            gA = torch.zeros((dims,) * 3, device='cuda')
            MASL.MA(dA, gA, box, 'PCS')           # queued on `side`, no wait on return
        gB = np.zeros((dims,) * 3, np.float32)
        MASL.MA(posB, gB, box, 'CIC')             # own stream of the same context
        side.synchronize()
        assert_cells(gA.cpu().numpy(), refA, what='side-stream PCS round %d' % it,
                     nadds=contrib_counts(posA, dims, box, 'PCS'))
        assert_cells(gB, refB, what='host CIC round %d' % it,
                     nadds=contrib_counts(posB, dims, box, 'CIC'))


def test_serving_waits_for_submitter_stream(oracle):
    """Positions written by a long kernel on the default stream, submitted at once:
    the served spectrum equals the drop-in API's on the finished positions, and
    the oracle's on the drop-in grid's delta (a read of half-written positions
    would give neither)."""
    import torch
    from pylians3_amd import MAS_library as MASL, Pk_library as PKL
    from pylians3_amd.serving import Request, Server
    box, dims, n = 400.0, 64, 200_000
    rng = np.random.default_rng(12)
    host = (rng.random((n, 3)) * box).astype(np.float32)
    with Server(slots=2) as srv:
        src = torch.from_numpy(host).cuda()
        torch.cuda.synchronize()
        # a long chain of default-stream work ending in the real positions
        junk = torch.rand((4096, 4096), device='cuda')
        for _ in range(20):
            junk = junk @ junk
            junk = junk / junk.abs().max()
        pos = torch.empty_like(src)
        pos.copy_(src + 0.0 * junk[0, 0])
        res = srv.submit(Request(pos, dims, box, 'TSC', 2)).result(timeout=120)
    grid = np.zeros((dims,) * 3, np.float32)
    MASL.MA(host, grid, box, 'TSC')
    delta = grid / np.float32(n / dims ** 3) - np.float32(1.0)
    ref = PKL.Pk(delta, box, axis=2, MAS='TSC', verbose=False)
    np.testing.assert_allclose(res.Pk, ref.Pk, rtol=1e-6, atol=1e-9 * np.abs(ref.Pk[:, 0]).max())
    ora = oracle.pk(delta, box, axis=2, MAS='TSC')
    from conftest import assert_bins
    np.testing.assert_array_equal(res.Nmodes3D, ora.Nmodes3D)
    for l in range(3):
        assert_bins(res.Pk[:, l], ora.Pk[:, l], ora.Pk[:, 0], what='served Pk l=%d' % (2 * l))
