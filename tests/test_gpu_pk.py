"""GPU parity of the power-spectrum path (rocFFT + the fused binning kernel).

Against the oracle (float32 FFT + the reference mode loop of
Pk_library.pyx:311-378 / :623-732 + its normalisation): Nmodes and k bins
bit-exact, Pk per bin within 1e-5 * max(|P_ref|, |P0_ref|) (the monopole is the
floor for the l = 2, 4 multipoles, whose bins can cross zero).  The binning
kernel alone is checked on exactly the oracle's Fourier data.
"""
import numpy as np
import pytest

from conftest import assert_bins

pytestmark = pytest.mark.gpu


def _field(dims, box, seed, n=None, mas='CIC'):
    """delta = rho/mean - 1 of random particles assigned with the oracle."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    n = n or 4 * dims ** 3
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    g = np.zeros((dims,) * 3, np.float32)
    O.ma_f32(pos, g, box, mas)
    return g / np.float32(g.mean()) - np.float32(1.0)


def _compare(r, ref, what):
    np.testing.assert_array_equal(r.Nmodes3D, ref.Nmodes3D)
    np.testing.assert_array_equal(r.Nmodes2D, ref.Nmodes2D)
    np.testing.assert_array_equal(r.Nmodes1D, ref.Nmodes1D)
    np.testing.assert_array_equal(r.k1D, ref.k1D)
    np.testing.assert_array_equal(r.kpar, ref.kpar)
    np.testing.assert_array_equal(r.kper, ref.kper)
    np.testing.assert_allclose(r.k3D, ref.k3D, rtol=1e-13)
    P0 = ref.Pk[:, 0]
    assert_bins(r.Pk[:, 0], P0, P0, what=what + ' l=0')
    assert_bins(r.Pk[:, 1], ref.Pk[:, 1], P0, what=what + ' l=2')
    assert_bins(r.Pk[:, 2], ref.Pk[:, 2], P0, what=what + ' l=4')
    assert_bins(r.Pk1D, ref.Pk1D, ref.Pk1D, what=what + ' 1D')
    # 2D bin (0,0) holds the DC mode only (rounding noise of a mean-subtracted field)
    fl2 = np.full_like(ref.Pk2D, np.median(np.abs(ref.Pk2D)))
    assert_bins(r.Pk2D, ref.Pk2D, fl2, what=what + ' 2D')


@pytest.mark.parametrize('dims', [16, 32, 33, 64])
@pytest.mark.parametrize('axis', [0, 1, 2])
def test_pk_vs_oracle(oracle, dims, axis):
    print('case pk', dims, axis, flush=True)
    from pylians3_amd import Pk_library as PKL
    box = 500.0
    delta = _field(dims, box, dims + axis)
    for mas in ('NGP', 'CIC', 'TSC', 'PCS', 'None'):
        r = PKL.Pk(delta, box, axis=axis, MAS=mas, verbose=False)
        ref = oracle.pk(delta, box, axis=axis, MAS=mas)
        _compare(r, ref, 'Pk d=%d axis=%d %s' % (dims, axis, mas))
        assert_bins(r.Pkphase, ref.Pkphase, ref.Pkphase, rtol=1e-5, what='phase')


@pytest.mark.parametrize('axis', [0, 1, 2])
@pytest.mark.parametrize('F', [1, 2])
def test_binning_alone_on_oracle_fourier_data(oracle, axis, F):
    """Same complex64 input to both: only the binning arithmetic differs."""
    print('case binning', axis, F, flush=True)
    from pylians3_amd import Pk_library as PKL
    dims, box = 48, 300.0
    deltas = [_field(dims, box, 40 + f) for f in range(F)]
    dks = [oracle.fft_f32(d) for d in deltas]
    mas = ['TSC', 'CIC'][:F]
    r = PKL.Pk_from_fourier(dks, box, dims, axis=axis, MAS=mas)
    ref = oracle.raw_bins(dks, dims, axis, mas)
    fin = oracle.finalize_pk(ref, dims, box, list(range(r['Pk3D'].shape[2])))
    np.testing.assert_array_equal(r['Nmodes3D'], fin.Nmodes3D)
    np.testing.assert_array_equal(r['Nmodes2D'], fin.Nmodes2D)
    P0 = fin.Pk[:, 0]
    # bin sums are exact integers in units of 2^-78 of a bound on the field's power:
    # a cross spectrum that cancels to near zero is held to its floor sqrt(P_i P_j)
    fl0 = np.abs(P0).copy()
    if F == 2:
        fl0[:, 2] = np.sqrt(np.abs(P0[:, 0] * P0[:, 1]))
    np.testing.assert_array_less(np.abs(r['Pk3D'][:, 0] - P0), 1e-12 * fl0 + 1e-300)
    np.testing.assert_allclose(r['Pk1D'], fin.Pk1D, rtol=1e-12, atol=1e-13 * np.abs(fin.Pk1D).max())
    np.testing.assert_allclose(r['Pk2D'], fin.Pk2D, rtol=1e-12, atol=1e-12 * np.abs(fin.Pk2D).max())
    # Legendre weights are evaluated in float32 on the GPU.  Floors: the monopole of
    # each auto-spectrum; sqrt(P0_i P0_j) for the cross spectrum of fields i, j
    S = r['Pk3D'].shape[2]
    autos = np.abs(P0[:, :F])
    floor = np.empty_like(P0)
    floor[:, :F] = autos
    if F == 2:
        floor[:, 2] = np.sqrt(autos[:, 0] * autos[:, 1])
    assert S == floor.shape[1]
    for l in (1, 2):
        assert_bins(r['Pk3D'][:, l], fin.Pk[:, l], floor, rtol=2e-6, what='l=%d' % (2 * l))
    if F == 1:
        phase = (ref['Pkphase'][1:] / fin.Nmodes3D) * (box / dims ** 2) ** 3
        assert_bins(r['Pkphase'], phase, phase, rtol=1e-5, what='phase')


@pytest.mark.parametrize('axis', [0, 2])
def test_xpk_vs_oracle(oracle, axis):
    print('case xpk', axis, flush=True)
    from pylians3_amd import Pk_library as PKL
    dims, box = 32, 400.0
    d1 = _field(dims, box, 70)
    d2 = _field(dims, box, 71, mas='TSC')
    d3 = _field(dims, box, 72)
    for fields, mas in (([d1, d2], ['CIC', 'TSC']), ([d1, d2, d3], ['CIC', 'TSC', 'PCS'])):
        r = PKL.XPk(fields, box, axis=axis, MAS=mas)
        ref = oracle.xpk(fields, box, axis=axis, MAS=mas)
        np.testing.assert_array_equal(r.Nmodes3D, ref.Nmodes3D)
        for f in range(len(fields)):
            P0 = ref.Pk[:, 0, f]
            for l in range(3):
                assert_bins(r.Pk[:, l, f], ref.Pk[:, l, f], P0, what='auto %d l%d' % (f, l))
            assert_bins(r.Pk1D[:, f], ref.Pk1D[:, f], ref.Pk1D[:, f], what='auto 1D')
        for x in range(r.XPk.shape[2]):
            # cross spectra cross zero: floor by the geometric mean of the autos
            i, j = [(a, b) for a in range(len(fields)) for b in range(a + 1, len(fields))][x]
            fl = np.sqrt(np.abs(ref.Pk[:, 0, i] * ref.Pk[:, 0, j]))
            for l in range(3):
                assert_bins(r.XPk[:, l, x], ref.XPk[:, l, x], fl, what='cross %d l%d' % (x, l))
            fl1 = np.sqrt(np.abs(ref.Pk1D[:, i] * ref.Pk1D[:, j]))
            assert_bins(r.PkX1D[:, x], ref.PkX1D[:, x], fl1, what='cross 1D')
            fl2 = np.sqrt(np.abs(ref.Pk2D[:, i] * ref.Pk2D[:, j]))
            fl2 = np.maximum(fl2, np.median(fl2))
            assert_bins(r.PkX2D[:, x], ref.PkX2D[:, x], fl2, what='cross 2D')


def test_fft_vs_direct(oracle):
    from pylians3_amd import Pk_library as PKL
    rng = np.random.default_rng(3)
    for d in (6, 16, 64):
        x = rng.standard_normal((d, d, d)).astype(np.float32)
        got = PKL.FFT3Dr_f(x, 1)
        ref = np.fft.rfftn(x.astype(np.float64))
        assert got.shape == ref.shape and got.dtype == np.complex64
        np.testing.assert_allclose(got, ref, rtol=0, atol=3e-6 * np.abs(ref).max())
        np.testing.assert_array_equal(x, x)   # input untouched


def test_plane_wave_on_gpu(oracle):
    from pylians3_amd import Pk_library as PKL
    d, L, A = 32, 100.0, 1.5
    i = np.arange(d)
    X, Y, Z = np.meshgrid(i, i, i, indexing='ij')
    x = (A * np.cos(2 * np.pi * (3 * X + 0 * Y + 4 * Z) / d)).astype(np.float32)
    r = PKL.Pk(x, L, axis=2, MAS='None', verbose=False)
    fact = (L / d ** 2) ** 3
    P = r.Pk[4, 0]          # |k| = 5
    np.testing.assert_allclose(P, (A * d ** 3 / 2) ** 2 * fact / r.Nmodes3D[4], rtol=1e-5)
    np.testing.assert_allclose(r.Pk[4, 1], 5 * P * (3 * 0.64 - 1) / 2, rtol=1e-5)
    others = np.delete(r.Pk[:, 0], 4)
    assert np.all(np.abs(others) < 1e-8 * P)


def test_pk_device_tensor_and_input_preserved(oracle):
    import torch
    from pylians3_amd import Pk_library as PKL
    dims, box = 32, 200.0
    delta = _field(dims, box, 90)
    t = torch.from_numpy(delta).cuda()
    before = t.clone()
    r = PKL.Pk(t, box, axis=1, MAS='CIC', verbose=False)
    assert torch.equal(t, before)
    ref = oracle.pk(delta, box, axis=1, MAS='CIC')
    _compare(r, ref, 'device')


def test_pk_deterministic(oracle):
    from pylians3_amd import Pk_library as PKL
    delta = _field(64, 500.0, 91)
    a = PKL.Pk(delta, 500.0, axis=0, MAS='CIC', verbose=False)
    b = PKL.Pk(delta, 500.0, axis=0, MAS='CIC', verbose=False)
    np.testing.assert_array_equal(a.Pk, b.Pk)
    np.testing.assert_array_equal(a.Pk2D, b.Pk2D)


def _pk_hash(arrs):
    import hashlib
    return hashlib.sha1(b''.join(np.ascontiguousarray(v).tobytes() for _, v in sorted(arrs.items()))).hexdigest()


@pytest.mark.parametrize('axis', [0, 2])
def test_pk_bitwise_reproducible_large(axis):
    """Every output of Pk (with the phase) and of XPk on a 512^3 field is the same
    bits on every call, with other work on the device between the calls: the bin
    sums are exact integers, so the order in which thousands of waves add into them
    does not matter (with double sums the last bits changed from call to call)."""
    import ctypes
    import torch
    from pylians3_amd import _native as N
    print('case bitwise Pk 512 axis', axis, flush=True)
    L = N.lib()
    d = 512
    g = torch.Generator(device='cuda').manual_seed(11)
    x = torch.randn((d, d, d), generator=g, device='cuda')
    delta = (torch.exp(x) / float(np.exp(0.5)) - 1.0).contiguous()
    del x
    ctx = N.Context(0)
    try:
        for nfl in (1, 2):
            hs = []
            out, arrs, _ = N.alloc_pk_out(d, nfl)
            for rep in range(3):
                ptrs = (ctypes.c_void_p * nfl)(*([delta.data_ptr()] * nfl))
                with N.torch_stream(ctx, True):
                    N.check(L.pyl_pk(ctx.handle, ptrs, nfl, (ctypes.c_int * nfl)(*([3] * nfl)), d,
                                     ctypes.c_double(1000.0), axis, ctypes.c_float(1.0), ctypes.c_float(0.0),
                                     N.PTR_DEVICE, ctypes.byref(out)))
                hs.append(_pk_hash(arrs))
                junk = torch.rand((2048, 2048), device='cuda')
                junk = junk @ junk                     # other work between the calls
                torch.cuda.synchronize()
            assert len(set(hs)) == 1, (nfl, hs)
            assert np.all(np.isfinite(arrs['Pk3D']))
    finally:
        ctx.close()


def test_pk_batch_bitwise_equals_single():
    """A realisation binned inside a batch gives the same bits as alone (each
    realisation has its own scale for the exact sums)."""
    import ctypes
    import torch
    from pylians3_amd import _native as N
    L = N.lib()
    d, nreal = 64, 5
    g = torch.Generator(device='cuda').manual_seed(12)
    x = torch.randn((nreal, d, d, d), generator=g, device='cuda')
    # very different amplitudes per realisation: different scales
    amp = torch.tensor([1e-3, 1.0, 30.0, 1e3, 1e-1], device='cuda').view(nreal, 1, 1, 1)
    deltas = (x * amp).contiguous()
    ctx = N.Context(0)
    try:
        outs, keep = [], []
        for r in range(nreal):
            o, a, _ = N.alloc_pk_out(d, 1)
            outs.append(o)
            keep.append(a)
        arr = (N.PkOut * nreal)(*outs)
        with N.torch_stream(ctx, True):
            N.check(L.pyl_pk_batch(ctx.handle, (ctypes.c_void_p * 1)(deltas.data_ptr()), nreal, 1,
                                   (ctypes.c_int * 1)(2), d, ctypes.c_double(300.0), 1, None, None,
                                   N.PTR_DEVICE, arr))
        for r in range(nreal):
            o, a, _ = N.alloc_pk_out(d, 1)
            one = deltas[r].contiguous()
            with N.torch_stream(ctx, True):
                N.check(L.pyl_pk(ctx.handle, (ctypes.c_void_p * 1)(one.data_ptr()), 1,
                                 (ctypes.c_int * 1)(2), d, ctypes.c_double(300.0), 1, ctypes.c_float(1.0),
                                 ctypes.c_float(0.0), N.PTR_DEVICE, ctypes.byref(o)))
            for k in ('Pk3D', 'Pk2D', 'Pk1D', 'Pkphase', 'k3D'):
                np.testing.assert_array_equal(keep[r][k], a[k], err_msg='r=%d %s' % (r, k))
    finally:
        ctx.close()


@pytest.mark.parametrize('dims', [16, 64, 128, 256, 512])
def test_fft_power_of_two_vs_scipy(oracle, dims):
    # the own z/y/x passes (power-of-two sides) against scipy's pocketfft in
    # float64: max error within a few float32 ulps of the largest mode, and the
    # float32 oracle FFT is no closer
    import torch
    from pylians3_amd import Pk_library as PKL
    print('case fft', dims, flush=True)
    rng = np.random.default_rng(dims)
    x = rng.standard_normal((dims,) * 3).astype(np.float32)
    ref = oracle.fft_f64(x)
    got = PKL.FFT3Dr_f(torch.from_numpy(x).cuda(), 1).cpu().numpy()
    scale = np.sqrt(np.mean(np.abs(ref) ** 2))
    err = np.abs(got - ref).max() / scale
    err32 = np.abs(oracle.fft_f32(x) - ref).max() / scale
    assert err < max(3 * err32, 1e-6), (err, err32)
    # a plane wave lands in exactly one mode with amplitude d^3/2
    i = np.arange(dims)
    X, Y, Z = np.meshgrid(i, i, i, indexing='ij')
    pw = np.cos(2 * np.pi * (3 * X + 1 * Y + 5 * Z) / dims).astype(np.float32)
    g = PKL.FFT3Dr_f(pw, 1)
    assert abs(g[3, 1, 5] - dims ** 3 / 2) < 1e-4 * dims ** 3
    g[3, 1, 5] = 0
    assert np.abs(g).max() < 1e-4 * dims ** 3


def test_pkphase_closed_form_angle_sweep():
    """Pkphase on exact Fourier data: one mode per 3D bin, at angles swept over
    (-pi, pi] and magnitudes over 60 decades; each bin holds atan2(Re, |delta|)^2
    of its mode (Pk_library.pyx:358,377) / Nmodes3D * (L/d^2)^3 (:418)."""
    from oracle import oracle as O
    from pylians3_amd import Pk_library as PKL
    dims, L = 64, 300.0
    m = dims // 2
    kmax = int(np.sqrt(3 * m * m))
    spots = {}
    for kx in range(-m + 1, m + 1):
        for ky in range(-m + 1, m + 1):
            K = int(np.floor(np.sqrt(kx * kx + ky * ky + 1)))
            spots.setdefault(K, (kx % dims, ky % dims, 1))
    bins = sorted(k for k in spots if 1 <= k <= kmax)
    _, _, N3 = O.brute_mode_counts(dims, 0)
    rng = np.random.default_rng(5)
    for rep in range(8):
        theta = np.linspace(-np.pi, np.pi, len(bins), endpoint=False) + rep * 0.37
        mag = 10.0 ** rng.uniform(-30, 30, len(bins))
        dk = np.zeros((dims, dims, m + 1), np.complex64)
        raw = np.zeros(kmax + 1)
        for K, t, a in zip(bins, theta, mag):
            v = np.complex64(a * np.exp(1j * t))
            dk[spots[K]] = v
            re, im = float(v.real), float(v.imag)
            raw[K] = np.arctan2(re, np.sqrt(re * re + im * im)) ** 2
        exp = raw[1:] / N3[1:] * (L / dims ** 2) ** 3
        r = PKL.Pk_from_fourier([dk], L, dims, axis=0, MAS=['None'])
        got = r['Pkphase']
        np.testing.assert_allclose(got, exp, rtol=2e-6, atol=1e-12 * exp.max(),
                                   err_msg='angle sweep %d' % rep)
