"""Pins the 2D-spectrum and correct_MAS parts of the oracle.

Pk_plane / XPk_plane (Pk_library.pyx:435-511, :1093-1220): mode counts against
check_number_modes_2D's closed form (Pk_library.pyx:102-114) and a brute-force
Hermitian enumeration (the reference's own test_number_modes_2D method,
Pk_library.pyx:1602-1653); single 2D plane waves with a closed-form bin value;
the 2D MAS window; XPk_plane([d, d]) == Pk_plane and XPk_plane([d, -d]) == -Pk.
XPk_2D (Pk_library.pyx:1754-1865): its bins equal the 2D bins of the 3D loop,
its DC bin is kept and the cross spectrum of a field with itself is its auto.
correct_MAS (Pk_library.pyx:1882-1939): the unnormalised inverse (dims^3 times
the field), the window divided out of a plane wave off the kz = 0 plane, and
(1 + c)/2 on a plane wave in it (the Hermitian part read by the real inverse
transform).
"""
import numpy as np
import pytest


def plane_wave_2d(g, n, A=1.0):
    i = np.arange(g)
    X, Y = np.meshgrid(i, i, indexing='ij')
    return A * np.cos(2 * np.pi * (n[0] * X + n[1] * Y) / g)


@pytest.mark.parametrize('grid', [4, 5, 6, 7, 8, 16, 33])
def test_plane_mode_counts(oracle, grid):
    d = np.random.default_rng(grid).standard_normal((grid, grid)).astype(np.float32)
    r = oracle.pk_plane(d, 100.0, 'CIC')
    np.testing.assert_array_equal(r.raw['N'], oracle.brute_mode_counts_2D(grid))
    assert r.raw['N'].sum() == oracle.expected_modes_2D(grid)


@pytest.mark.parametrize('n', [(3, 0), (0, 5), (2, -3), (4, 4), (-5, 1)])
def test_plane_wave_closed_form(oracle, n):
    g, L, A = 16, 300.0, 0.6
    x = plane_wave_2d(g, n, A)
    r = oracle.pk_plane(x, L, 'None', f64=True)
    k2 = n[0] ** 2 + n[1] ** 2
    b = int(np.floor(np.sqrt(k2)))
    # one mode of amplitude A g^2 / 2 in the half plane (its partner is redundant)
    expect = np.zeros_like(r.Pk)
    expect[b - 1] = (A * g * g / 2) ** 2 / r.Nmodes[b - 1] * (L / g ** 2) ** 2
    np.testing.assert_allclose(r.Pk, expect, rtol=1e-12, atol=1e-12 * expect.max())
    # mean |k| of the bin is from the mode counts only
    assert np.all(r.k > 0)


def test_plane_mas_window(oracle):
    g, L = 16, 100.0
    n = (3, 2)
    x = plane_wave_2d(g, n)
    r0 = oracle.pk_plane(x, L, 'None', f64=True)
    for mas, p in (('NGP', 1), ('CIC', 2), ('TSC', 3), ('PCS', 4)):
        r = oracle.pk_plane(x, L, mas, f64=True)
        c = 1.0
        for ni in n:
            s = np.pi * ni / g
            c *= (s / np.sin(s)) ** p
        b = int(np.floor(np.hypot(*n)))
        np.testing.assert_allclose(r.Pk[b - 1], r0.Pk[b - 1] * c * c, rtol=1e-12)


def test_xpk_plane_identities(oracle):
    g, L = 32, 150.0
    rng = np.random.default_rng(3)
    d = rng.standard_normal((g, g)).astype(np.float32)
    a = oracle.pk_plane(d, L, 'CIC')
    x = oracle.xpk_plane(d, d, L, 'CIC', 'CIC')
    np.testing.assert_array_equal(x.Nmodes, a.Nmodes)
    np.testing.assert_allclose(x.Pk[:, 0], a.Pk, rtol=1e-14)
    np.testing.assert_allclose(x.Pk[:, 1], a.Pk, rtol=1e-14)
    np.testing.assert_allclose(x.XPk, a.Pk, rtol=1e-14)
    np.testing.assert_allclose(x.r, 1.0, rtol=1e-14)
    y = oracle.xpk_plane(d, -d, L, 'CIC', 'CIC')
    np.testing.assert_allclose(y.XPk, -a.Pk, rtol=1e-14)
    # MAS of the second map only changes the second auto and the cross by c and sqrt(c)
    z = oracle.xpk_plane(d, d, L, 'CIC', 'None')
    assert np.all(z.Pk[:, 1] <= z.Pk[:, 0] * (1 + 1e-12))


def test_xpk_2d_bins(oracle):
    dims, L = 16, 200.0
    rng = np.random.default_rng(7)
    d1 = rng.standard_normal((dims,) * 3).astype(np.float32)
    d2 = rng.standard_normal((dims,) * 3).astype(np.float32)
    kpar, kper, P1, P2, PX, N = oracle.xpk_2d(d1, d2, L, axis=1, MAS1='CIC', MAS2='TSC')
    x = oracle.xpk([d1, d2], L, axis=1, MAS=['CIC', 'TSC'])
    np.testing.assert_array_equal(N, x.Nmodes2D)
    np.testing.assert_array_equal(kpar, x.kpar)
    np.testing.assert_array_equal(kper, x.kper)
    ok = N > 0
    np.testing.assert_allclose(P1[ok], x.Pk2D[ok, 0], rtol=1e-14)
    np.testing.assert_allclose(P2[ok], x.Pk2D[ok, 1], rtol=1e-14)
    np.testing.assert_allclose(PX[ok], x.PkX2D[ok, 0], rtol=1e-14)
    _, _, Q1, Q2, QX, _ = oracle.xpk_2d(d1, d1, L, axis=1, MAS1='CIC', MAS2='CIC')
    np.testing.assert_allclose(QX[ok], Q1[ok], rtol=1e-14)
    N1, N2, N3 = oracle.brute_mode_counts(dims, 1)
    np.testing.assert_array_equal(N, N2)


def test_correct_mas_none_is_scaled_identity(oracle):
    d = 12
    x = np.random.default_rng(0).standard_normal((d, d, d))
    y = oracle.correct_mas(x, 100.0, 'None', f64=True)
    np.testing.assert_allclose(y, x * d ** 3, rtol=0, atol=1e-10 * d ** 3)


@pytest.mark.parametrize('mas,p', [('NGP', 1), ('CIC', 2), ('TSC', 3), ('PCS', 4)])
def test_correct_mas_plane_waves(oracle, mas, p):
    d = 16
    i = np.arange(d)
    X, Y, Z = np.meshgrid(i, i, i, indexing='ij')

    def window(n):
        c = 1.0
        for ni in n:
            if ni:
                s = np.pi * ni / d
                c *= (s / np.sin(s)) ** p
        return c

    # off the kz = 0 plane: both members of the pair are reached through the
    # independent half, so the window is divided out exactly
    n = (2, -3, 5)
    x = np.cos(2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / d)
    y = oracle.correct_mas(x, 100.0, mas, f64=True)
    np.testing.assert_allclose(y, x * d ** 3 * window(n), rtol=0, atol=1e-9 * d ** 3)
    # in the kz = 0 plane the redundant partner keeps the raw amplitude: the real
    # inverse transform sees (1 + c)/2
    n = (3, 2, 0)
    x = np.cos(2 * np.pi * (n[0] * X + n[1] * Y + n[2] * Z) / d)
    y = oracle.correct_mas(x, 100.0, mas, f64=True)
    np.testing.assert_allclose(y, x * d ** 3 * (1 + window(n)) / 2, rtol=0, atol=1e-9 * d ** 3)
