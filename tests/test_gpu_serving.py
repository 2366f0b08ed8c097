"""Continuous-batching server: every request's result is the same alone and in the
mix (bitwise: the tiled scatter and the binning do not depend on batching or on
concurrent streams), admission respects the HBM budget, and fused batches form
when compatible requests wait together."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(seed, n=10):
    import torch
    from pylians3_amd.serving import Request
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        dims = int(rng.choice([32, 64]))
        npart = int(rng.integers(10000, 200000))
        mas = str(rng.choice(['NGP', 'CIC', 'TSC', 'PCS']))
        axis = int(rng.integers(0, 3))
        g = torch.Generator(device='cuda').manual_seed(1000 + seed * 100 + i)
        pos = torch.rand((npart, 3), generator=g, device='cuda') * 300.0
        if i % 3 == 0:
            pos2 = torch.rand((npart // 2, 3), generator=g, device='cuda') * 300.0
            W2 = torch.rand(npart // 2, generator=g, device='cuda') + 0.5
            reqs.append(Request(pos, dims, 300.0, mas, axis, pos2=pos2, W2=W2))
        else:
            reqs.append(Request(pos, dims, 300.0, mas, axis))
    return reqs


def _same(a, b):
    np.testing.assert_array_equal(a.Pk, b.Pk)
    np.testing.assert_array_equal(a.Pk2D, b.Pk2D)
    np.testing.assert_array_equal(a.Pk1D, b.Pk1D)
    if hasattr(a, 'XPk'):
        np.testing.assert_array_equal(a.XPk, b.XPk)


def test_results_same_alone_and_in_mix():
    from pylians3_amd.serving import Server
    reqs = _requests(1)
    alone = []
    with Server(slots=1, max_batch=1) as srv:
        for r in reqs:
            alone.append(srv.submit(r).result(timeout=120))
    with Server(slots=3, max_batch=8) as srv:
        futs = [srv.submit(r) for r in reqs]
        mixed = [f.result(timeout=120) for f in futs]
        batches = [m.batch for m in mixed]
    for a, m in zip(alone, mixed):
        _same(a, m)
    assert max(batches) >= 1


def test_matches_drop_in_api(oracle):
    import torch
    from pylians3_amd import MAS_library as MASL, Pk_library as PKL
    from pylians3_amd.serving import Request, Server
    g = torch.Generator(device='cuda').manual_seed(7)
    pos = torch.rand((100000, 3), generator=g, device='cuda') * 500.0
    with Server(slots=2) as srv:
        res = srv.submit(Request(pos, 64, 500.0, 'TSC', 1)).result(timeout=120)
    grid = np.zeros((64,) * 3, np.float32)
    MASL.MA(pos.cpu().numpy(), grid, 500.0, 'TSC')
    delta = grid / np.float32(100000 / 64 ** 3) - np.float32(1.0)
    ref = PKL.Pk(delta, 500.0, axis=1, MAS='TSC', verbose=False)
    np.testing.assert_allclose(res.Pk, ref.Pk, rtol=1e-6, atol=1e-9 * np.abs(ref.Pk[:, 0]).max())
    np.testing.assert_array_equal(res.Nmodes3D, ref.Nmodes3D)
    # and against the oracle on the same float32 delta
    from conftest import assert_bins
    ora = oracle.pk(delta, 500.0, axis=1, MAS='TSC')
    np.testing.assert_array_equal(res.Nmodes3D, ora.Nmodes3D)
    for l in range(3):
        assert_bins(res.Pk[:, l], ora.Pk[:, l], ora.Pk[:, 0], what='served vs oracle l=%d' % (2 * l))


def test_budget_admission_and_rejection():
    import torch
    from pylians3_amd import _native as N
    from pylians3_amd.serving import Request, Server
    reqs = _requests(2, n=6)
    need = max(r.hbm_bytes() for r in reqs)
    with Server(hbm_budget=need, slots=4) as srv:     # at most one request in flight
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=120) for f in futs]
        assert len(res) == len(reqs)
        big = Request(torch.zeros((10, 3), device='cuda'), 1024, 1.0)
        with pytest.raises(N.BudgetError):
            srv.submit(big)


def test_warm_runs_every_slot_and_keeps_results():
    # warm() runs a request once on every slot (growing each context's scratch); the
    # results served afterwards are the same bits as on an unwarmed server
    from pylians3_amd.serving import Server
    reqs = _requests(3, n=4)
    with Server(slots=1, max_batch=1) as srv:
        cold = [srv.submit(r).result(timeout=120) for r in reqs]
    with Server(slots=3, max_batch=1) as srv:
        srv.warm(max(reqs, key=lambda r: r.hbm_bytes()))
        warm = [srv.submit(r).result(timeout=120) for r in reqs]
    for a, b in zip(cold, warm):
        _same(a, b)
