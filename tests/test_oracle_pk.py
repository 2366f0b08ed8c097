"""Pins the power-spectrum oracle (oracle/pk_oracle.c + oracle/oracle.py).

Mode counts against the closed form of check_number_modes (Pk_library.pyx:87-99)
and a brute-force Hermitian enumeration of all dims^3 modes; the FFT against a
direct DFT, plane waves and Parseval; bin values against single plane waves
with closed-form amplitudes; the Legendre weights with modes on the line of
sight; the MAS window against x/sin(x) powers; XPk against auto-spectrum
identities; the float32 path against the float64 oracle.
"""
import numpy as np
import pytest


@pytest.mark.parametrize('dims', [4, 5, 6, 8, 9, 16])
@pytest.mark.parametrize('axis', [0, 1, 2])
def test_mode_counts_brute_force(oracle, dims, axis):
    x = np.random.default_rng(dims).standard_normal((dims,) * 3).astype(np.float32)
    raw = oracle.pk(x, 50.0, axis=axis).raw
    N1, N2, N3 = oracle.brute_mode_counts(dims, axis)
    np.testing.assert_array_equal(raw['N3D'], N3)
    np.testing.assert_array_equal(raw['N2D'], N2)
    np.testing.assert_array_equal(raw['N1D'], N1)
    assert raw['N3D'].sum() == oracle.expected_modes(dims)
    # k1D sums |k_par| over 1D modes
    np.testing.assert_array_equal(raw['k1D'], np.arange(dims // 2 + 1) * N1)


def test_fft_direct_dft_and_parseval(oracle):
    rng = np.random.default_rng(1)
    for d in (4, 5, 6):
        x = rng.standard_normal((d, d, d)).astype(np.float32)
        ref = oracle.direct_dft_r2c(x.astype(np.float64))
        got = oracle.fft_f32(x)
        assert got.dtype == np.complex64
        np.testing.assert_allclose(got, ref, rtol=0, atol=2e-5 * np.abs(ref).max())
    x = rng.standard_normal((16, 16, 16))
    X = np.fft.fftn(x)
    assert abs(np.sum(np.abs(X) ** 2) / 16 ** 3 - np.sum(x ** 2)) < 1e-9 * np.sum(x ** 2)


def plane_wave(d, n, A=1.0):
    i = np.arange(d)
    X, Y, Z = np.meshgrid(i, i, i, indexing='ij')
    return (A * np.cos(2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / d)).astype(np.float64)


@pytest.mark.parametrize('n,axis', [((3, 0, 0), 0), ((0, 2, 5), 2), ((1, 2, 2), 1),
                                    ((0, 0, 4), 2), ((2, -3, 1), 0)])
def test_plane_wave_closed_form(oracle, n, axis):
    d, L, A = 16, 200.0, 0.7
    x = plane_wave(d, n, A)
    r = oracle.pk(x.astype(np.float32), L, axis=axis, MAS='None', f64=True,
                  dk=oracle.fft_f64(x))
    k2 = n[0] ** 2 + n[1] ** 2 + n[2] ** 2
    kidx = int(np.floor(np.sqrt(k2)))
    fact = (L / d ** 2) ** 3
    N3 = r.Nmodes3D[kidx - 1]
    P0 = (A * d ** 3 / 2) ** 2 * fact / N3
    expect = np.zeros_like(r.Pk[:, 0])
    expect[kidx - 1] = P0
    np.testing.assert_allclose(r.Pk[:, 0], expect, rtol=1e-10, atol=1e-10 * P0)
    mu2 = n[axis] ** 2 / k2
    np.testing.assert_allclose(r.Pk[kidx - 1, 1], 5 * P0 * (3 * mu2 - 1) / 2, rtol=1e-10, atol=1e-10 * P0)
    np.testing.assert_allclose(r.Pk[kidx - 1, 2], 9 * P0 * (35 * mu2 ** 2 - 30 * mu2 + 3) / 8,
                               rtol=1e-10, atol=1e-10 * P0)
    np.testing.assert_allclose(r.k3D[kidx - 1], r.raw['k3D'][kidx] / N3 * 2 * np.pi / L, rtol=1e-14)
    # 2D: the single mode lands in its (k_par, k_per) bin, divided by that bin's count
    kmax_par = d // 2
    kpar = abs(n[axis])
    kper = int(np.floor(np.sqrt(k2 - n[axis] ** 2)))
    i2 = (kmax_par + 1) * kper + kpar
    P2 = np.zeros_like(r.Pk2D)
    P2[i2] = (A * d ** 3 / 2) ** 2 * fact / r.Nmodes2D[i2]
    np.testing.assert_allclose(r.Pk2D, P2, rtol=1e-10, atol=1e-10 * P2.max())
    # 1D (k <= middle): Pk1D = |dk|^2 fact * (pi kmaxper^2 / N) / (2 pi)^2
    if k2 <= (d // 2) ** 2:
        kF = 2 * np.pi / L
        kN = d // 2 * kF
        N1 = r.Nmodes1D[kpar - 1] if kpar > 0 else None
        if N1 is not None:
            kmaxper = np.sqrt(kN ** 2 - (kpar * kF) ** 2)
            v = (A * d ** 3 / 2) ** 2 * fact * (np.pi * kmaxper ** 2 / N1) / (2 * np.pi) ** 2
            np.testing.assert_allclose(r.Pk1D[kpar - 1], v, rtol=1e-10)


def test_line_of_sight_mode_legendre(oracle):
    # mu^2 = 1: Pk[:,1] = 5 Pk[:,0], Pk[:,2] = 9 Pk[:,0] in that bin
    for axis in (0, 1, 2):
        n = [0, 0, 0]
        n[axis] = 3
        x = plane_wave(16, n)
        r = oracle.pk(x.astype(np.float32), 100.0, axis=axis, MAS='None', f64=True,
                      dk=oracle.fft_f64(x))
        np.testing.assert_allclose(r.Pk[2, 1], 5 * r.Pk[2, 0], rtol=1e-12)
        np.testing.assert_allclose(r.Pk[2, 2], 9 * r.Pk[2, 0], rtol=1e-12)


def test_mas_correction_closed_form(oracle):
    for x in (0.1, 0.5, 1.2, -0.7, np.pi / 2):
        for p in (0, 1, 2, 3, 4):
            assert oracle.mas_correction(x, p) == pytest.approx((x / np.sin(x)) ** p, rel=1e-15)
    assert oracle.mas_correction(0.0, 3) == 1.0


def test_mas_deconvolution_single_mode(oracle):
    # Pk(MAS=m) / Pk(MAS='None') for one mode equals the squared float MAS factor
    d, n = 16, (2, 3, 1)
    x = plane_wave(d, n).astype(np.float32)
    dk = oracle.fft_f32(x)
    a = oracle.pk(x, 100.0, axis=2, MAS='None', dk=dk)
    for mas, p in (('CIC', 2), ('TSC', 3), ('PCS', 4)):
        b = oracle.pk(x, 100.0, axis=2, MAS=mas, dk=dk)
        f = np.float32(np.prod([oracle.mas_correction(np.pi / d * k, p) for k in n]))
        k = int(np.sqrt(sum(v * v for v in n)))
        np.testing.assert_allclose(b.Pk[k - 1, 0] / a.Pk[k - 1, 0], float(f) ** 2, rtol=2e-7)


def test_xpk_identities(oracle):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((16, 16, 16)).astype(np.float32)
    for axis in (0, 2):
        r = oracle.xpk([x, x], 300.0, axis=axis, MAS=['CIC', 'CIC'])
        np.testing.assert_allclose(r.XPk[:, :, 0], r.Pk[:, :, 0], rtol=1e-12)
        np.testing.assert_allclose(r.PkX2D[:, 0], r.Pk2D[:, 0], rtol=1e-12)
        np.testing.assert_allclose(r.PkX1D[:, 0], r.Pk1D[:, 0], rtol=1e-12)
        r2 = oracle.xpk([x, -x], 300.0, axis=axis, MAS=['CIC', 'CIC'])
        np.testing.assert_allclose(r2.XPk[:, :, 0], -r2.Pk[:, :, 0], rtol=1e-12)
        r3 = oracle.xpk([x, np.float32(2.0) * x], 300.0, axis=axis, MAS=['CIC', 'CIC'])
        np.testing.assert_allclose(r3.XPk[:, :, 0], 2 * r3.Pk[:, :, 0], rtol=1e-6)
        single = oracle.pk(x, 300.0, axis=axis, MAS='CIC')
        np.testing.assert_allclose(r.Pk[:, :, 0], single.Pk, rtol=1e-10)


def test_f32_path_close_to_fp64_oracle(oracle):
    rng = np.random.default_rng(6)
    pos = (rng.random((40000, 3)) * 500).astype(np.float32)
    g = np.zeros((32,) * 3, np.float32)
    oracle.ma_f32(pos, g, 500.0, 'CIC')
    delta = g / np.float32(g.mean()) - np.float32(1.0)
    a = oracle.pk(delta, 500.0, axis=0, MAS='CIC')
    g64 = oracle.ma_f64(pos, 32, 500.0, 'CIC')
    b = oracle.pk(g64 / g64.mean() - 1.0, 500.0, axis=0, MAS='CIC', f64=True)
    np.testing.assert_allclose(a.Pk[:, 0], b.Pk[:, 0], rtol=1e-4)
    np.testing.assert_array_equal(a.Nmodes3D, b.Nmodes3D)


@pytest.mark.parametrize('dims', [5, 8, 12, 16])
@pytest.mark.parametrize('axis', [0, 2])
def test_k3d_sums_brute_force(oracle, dims, axis):
    """raw k3D[k_index] += k (Pk_library.pyx:373): the sum of |k| over the bin's
    independent modes, from the brute-force enumeration."""
    x = np.random.default_rng(dims).standard_normal((dims,) * 3).astype(np.float32)
    r = oracle.pk(x, 80.0, axis=axis)
    np.testing.assert_allclose(r.raw['k3D'], oracle.brute_k3d_sums(dims, axis), rtol=1e-13)
    # finalized k3D is the mean |k| of the bin in units of kF (Pk_library.pyx:411)
    K = oracle.brute_k3d_sums(dims, axis)[1:]
    _, N2, N3 = oracle.brute_mode_counts(dims, axis)
    np.testing.assert_allclose(r.k3D, K / N3[1:] * (2 * np.pi / 80.0), rtol=1e-13)


@pytest.mark.parametrize('dims', [6, 9, 16])
@pytest.mark.parametrize('axis', [0, 1, 2])
def test_kpar_kper_bin_centres(oracle, dims, axis):
    """Every independent mode's |k_par| and |k_perp| fall inside the half-open
    bin [centre - kF/2, centre + kF/2) of the 2D bin it is counted in
    (index_2D = (kmax_par+1) k_per + k_par, Pk_library.pyx:367,399-400)."""
    L = 150.0
    kF = 2 * np.pi / L
    x = np.random.default_rng(3).standard_normal((dims,) * 3).astype(np.float32)
    r = oracle.pk(x, L, axis=axis)
    m = oracle.brute_modes(dims, axis)
    par = m['kpar'] * kF
    perp = np.sqrt((m['k2'] - m['kpar'] ** 2).astype(np.float64)) * kF
    cpar, cper = r.kpar[m['i2']], r.kper[m['i2']]
    assert np.all(par >= cpar - 0.5 * kF * (1 + 1e-12)) and np.all(par < cpar + 0.5 * kF)
    assert np.all(perp >= cper - 0.5 * kF * (1 + 1e-12)) and np.all(perp < cper + 0.5 * kF)
    # bins nobody reaches still have their centres: a full (kmax_par+1) x (kmax_per+1) table
    middle = dims // 2
    kmax_per = int(np.sqrt(2 * middle ** 2))
    assert r.kpar.shape == ((middle + 1) * (kmax_per + 1),)
    np.testing.assert_allclose(np.unique(r.kpar), (np.arange(middle + 1) + 0.5) * kF, rtol=1e-14)
    np.testing.assert_allclose(np.unique(r.kper), (np.arange(kmax_per + 1) + 0.5) * kF, rtol=1e-14)


def phase_field(dims, modes):
    """Half-complex field that is zero except at the given (kxx, kyy, kzz) -> value
    entries (0 < kzz < middle, so each is visited once by the mode loop)."""
    dk = np.zeros((dims, dims, dims // 2 + 1), np.complex64)
    for (i, j, l), v in modes.items():
        assert 0 < l < dims // 2
        dk[i, j, l] = v
    return dk


PHASE_MODES = {(1, 2, 3): 3 - 4j, (5, 0, 1): -1 + 0j, (0, 0, 7): 2j, (14, 9, 4): -2.5 + 0.5j}


def expected_phase(dims, L, modes):
    """Pkphase closed form: phase = atan2(Re, |delta_k|) per mode (Pk_library.pyx:358),
    phase^2 summed per 3D bin (:377), then / Nmodes3D * (L/d^2)^3 (:418).  A zero
    mode has atan2(0, 0) = 0."""
    kmax = int(np.sqrt(3 * (dims // 2) ** 2))
    raw = np.zeros(kmax + 1)
    for (i, j, l), v in modes.items():
        k = [a - dims if a > dims // 2 else a for a in (i, j, l)]
        kidx = int(np.floor(np.sqrt(sum(a * a for a in k))))
        raw[kidx] += np.arctan2(v.real, abs(v)) ** 2
    _, _, N3 = oracle_counts(dims)
    return raw[1:] / N3[1:] * (L / dims ** 2) ** 3


def oracle_counts(dims):
    from oracle import oracle as O
    return O.brute_mode_counts(dims, 0)


def test_pkphase_closed_form(oracle):
    dims, L = 16, 120.0
    dk = phase_field(dims, PHASE_MODES)
    r = oracle.pk(np.zeros((dims,) * 3, np.float32), L, axis=0, MAS='None', dk=dk)
    exp = expected_phase(dims, L, PHASE_MODES)
    np.testing.assert_allclose(r.Pkphase, exp, rtol=1e-15, atol=0)
    assert np.count_nonzero(r.Pkphase) == len({int(np.sqrt(sum((a - dims if a > dims // 2 else a) ** 2
                                                               for a in m)))
                                               for m in PHASE_MODES if PHASE_MODES[m].real != 0})


@pytest.mark.parametrize('F', [1, 2])
def test_threaded_loop_equals_serial(oracle, F):
    """The threaded mode loop (per-thread sums, added in thread order) gives the
    serial loop's counts exactly and its sums to double rounding, run after run."""
    rng = np.random.default_rng(9)
    for dims in (7, 16, 34):
        dks = [oracle.fft_f32(rng.standard_normal((dims,) * 3).astype(np.float32)) for _ in range(F)]
        for axis in (0, 2):
            a = oracle.raw_bins(dks, dims, axis, ['TSC', 'CIC'][:F])
            b = oracle.raw_bins(dks, dims, axis, ['TSC', 'CIC'][:F], threads=5)
            c = oracle.raw_bins(dks, dims, axis, ['TSC', 'CIC'][:F], threads=5)
            for k in a:
                np.testing.assert_array_equal(b[k], c[k])
                np.testing.assert_allclose(b[k], a[k], rtol=1e-13, atol=1e-13 * np.abs(a[k]).max())
            for k in ('N1D', 'N2D', 'N3D', 'k1D'):
                np.testing.assert_array_equal(a[k], b[k])


def test_same_delta_criterion_accepts_float32_ffts_and_rejects_mistakes(oracle):
    """The per-bin criterion the GPU spectra are held to on large clustered fields
    (conftest.assert_bins_same_delta): a second, independent float32 transform (a
    direct DFT by complex64 matrix products) of a steep-spectrum field passes it; a
    bin off by 1e-4, a bin swapped with its neighbour and a dropped Nyquist plane
    each fail it."""
    from conftest import assert_bins_same_delta
    d, box = 64, 500.0
    rng = np.random.default_rng(5)
    # a steep (k^-3) Gaussian field, exponentiated: bins span many decades of power
    kx = np.fft.fftfreq(d)[:, None, None]
    ky = np.fft.fftfreq(d)[None, :, None]
    kz = np.fft.rfftfreq(d)[None, None, :]
    k = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    amp = np.where(k > 0, k, 1.0) ** -1.5
    amp[0, 0, 0] = 0.0
    g = np.fft.irfftn(amp * (rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape)),
                      s=(d, d, d), axes=(0, 1, 2))
    g = np.exp(1.5 * g / g.std())
    delta = (g / g.mean() - 1.0).astype(np.float32)
    p32 = oracle.pk(delta, box, axis=0, MAS='CIC')
    p64 = oracle.pk(delta, box, axis=0, MAS='CIC', f64=True)
    n = np.arange(d)
    Wz = np.exp(-2j * np.pi * np.outer(n, np.arange(d // 2 + 1)) / d).astype(np.complex64)
    W = np.exp(-2j * np.pi * np.outer(n, n) / d).astype(np.complex64)
    dk_t = delta.astype(np.complex64) @ Wz
    dk_t = np.tensordot(W, dk_t, axes=(1, 0)).astype(np.complex64)
    dk_t = np.tensordot(W, dk_t, axes=(1, 1)).transpose(1, 0, 2).astype(np.complex64)
    assert dk_t.dtype == np.complex64
    np.testing.assert_allclose(dk_t, oracle.fft_f32(delta), rtol=0, atol=1e-4 * np.abs(dk_t).max())
    alt = oracle.pk(delta, box, axis=0, MAS='CIC', dk=dk_t)
    fl2 = np.full_like(p32.Pk2D, np.median(np.abs(p32.Pk2D)))
    cases = [(alt.Pk[:, l], p32.Pk[:, l], p64.Pk[:, l], p32.Pk[:, 0]) for l in range(3)]
    cases += [(alt.Pk2D, p32.Pk2D, p64.Pk2D, fl2), (alt.Pk1D, p32.Pk1D, p64.Pk1D, p32.Pk1D)]
    for got, a, b, fl in cases:
        assert_bins_same_delta(got, a, b, fl)
    P0 = alt.Pk[:, 0]
    off = P0.copy()
    off[3] *= 1 + 1e-4
    with pytest.raises(AssertionError):
        assert_bins_same_delta(off, p32.Pk[:, 0], p64.Pk[:, 0], p32.Pk[:, 0])
    swp = P0.copy()
    swp[[5, 6]] = swp[[6, 5]]
    with pytest.raises(AssertionError):
        assert_bins_same_delta(swp, p32.Pk[:, 0], p64.Pk[:, 0], p32.Pk[:, 0])
    dk_cut = dk_t.copy()
    dk_cut[d // 2] = 0
    cut = oracle.pk(delta, box, axis=0, MAS='CIC', dk=dk_cut)
    with pytest.raises(AssertionError):
        assert_bins_same_delta(cut.Pk2D, p32.Pk2D, p64.Pk2D, fl2)
