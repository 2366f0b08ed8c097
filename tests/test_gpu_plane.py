"""GPU parity of the 2D spectra (Pk_plane, XPk_plane, XPk_2D) and correct_MAS.

Against the oracle (float32 FFT + the reference mode loops of
Pk_library.pyx:466-497 / :1155-1203 / the 3D loop for XPk_2D, and the
correct_MAS mode scaling of :1910-1929 with an unnormalised inverse): Nmodes
bit-exact, mean k within 1e-13.  The 2D map spectra are checked against the
FP64 oracle: per bin within max(1e-5 * |P64|, 3x the float32 oracle's own
error) -- the maps are steep (red) spectra, whose low-power bins carry the
float32 FFT error of the whole map in any implementation; the cross spectrum
uses sqrt(P1 P2) as floor.  XPk_2D per bin within 1e-5 * max(|P_ref|, floor);
correct_MAS per cell within 1e-5 * max(|ref|, rms(ref)).
"""
import numpy as np
import pytest

from conftest import assert_bins, assert_bins_fp64

pytestmark = pytest.mark.gpu


def _map(grid, seed, kind='noise'):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((grid, grid))
    if kind == 'red':   # power concentrated at low k, like a projected density map
        import scipy.fft
        k = np.fft.fftfreq(grid)[:, None] ** 2 + np.fft.rfftfreq(grid)[None, :] ** 2
        X = scipy.fft.rfftn(x) / np.maximum(k, 1.0 / grid ** 2) ** 0.75
        x = scipy.fft.irfftn(X, s=x.shape)
    return x.astype(np.float32)


@pytest.mark.parametrize('grid', [16, 33, 64, 256])
@pytest.mark.parametrize('mas', ['CIC', 'PCS', 'None'])
def test_pk_plane_vs_oracle(oracle, grid, mas):
    print('case pk_plane', grid, mas, flush=True)
    from pylians3_amd import Pk_library as PKL
    d = _map(grid, grid, 'red')
    r = PKL.Pk_plane(d, 250.0, MAS=mas, verbose=False)
    ref = oracle.pk_plane(d, 250.0, mas)
    ref64 = oracle.pk_plane(d.astype(np.float64), 250.0, mas, f64=True)
    np.testing.assert_array_equal(r.Nmodes, ref.Nmodes)
    np.testing.assert_allclose(r.k, ref.k, rtol=1e-13)
    assert_bins_fp64(r.Pk, ref.Pk, ref64.Pk, what='Pk_plane')


@pytest.mark.parametrize('grid', [32, 65, 128])
def test_xpk_plane_vs_oracle(oracle, grid):
    print('case xpk_plane', grid, flush=True)
    from pylians3_amd import Pk_library as PKL
    d1 = _map(grid, 1, 'red')
    d2 = (0.5 * d1 + _map(grid, 2, 'red')).astype(np.float32)
    r = PKL.XPk_plane(d1, d2, 100.0, MAS1='CIC', MAS2='TSC')
    ref = oracle.xpk_plane(d1, d2, 100.0, 'CIC', 'TSC')
    r64 = oracle.xpk_plane(d1.astype(np.float64), d2.astype(np.float64), 100.0, 'CIC', 'TSC',
                           f64=True)
    np.testing.assert_array_equal(r.Nmodes, ref.Nmodes)
    np.testing.assert_allclose(r.k, ref.k, rtol=1e-13)
    assert_bins_fp64(r.Pk[:, 0], ref.Pk[:, 0], r64.Pk[:, 0], what='auto 1')
    assert_bins_fp64(r.Pk[:, 1], ref.Pk[:, 1], r64.Pk[:, 1], what='auto 2')
    fl = np.sqrt(r64.Pk[:, 0] * r64.Pk[:, 1])
    assert_bins_fp64(r.XPk, ref.XPk, r64.XPk, floor=fl, what='cross')
    np.testing.assert_allclose(r.r, r64.r, rtol=0, atol=3 * np.abs(ref.r - r64.r).max() + 2e-5)


def test_xpk_plane_device_and_identity(oracle):
    import torch
    from pylians3_amd import Pk_library as PKL
    d = _map(64, 9, 'red')
    t = torch.from_numpy(d).cuda()
    r = PKL.XPk_plane(t, t, 100.0, MAS1='CIC', MAS2='CIC')
    a = PKL.Pk_plane(d, 100.0, MAS='CIC', verbose=False)
    np.testing.assert_array_equal(r.Pk[:, 0], a.Pk)
    np.testing.assert_array_equal(r.XPk, a.Pk)


@pytest.mark.parametrize('axis', [0, 2])
def test_xpk_2d_vs_oracle(oracle, axis):
    print('case xpk_2d', axis, flush=True)
    from pylians3_amd import Pk_library as PKL
    rng = np.random.default_rng(axis)
    d1 = rng.standard_normal((32,) * 3).astype(np.float32)
    d2 = (d1 + rng.standard_normal((32,) * 3)).astype(np.float32)
    got = PKL.XPk_2D(d1, d2, 300.0, axis=axis, MAS1='CIC', MAS2='PCS')
    ref = oracle.xpk_2d(d1, d2, 300.0, axis=axis, MAS1='CIC', MAS2='PCS')
    for a, b in zip(got[:2], ref[:2]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got[5], ref[5])
    ok = ref[5] > 1
    med = np.median(np.abs(ref[2][ok]))
    assert_bins(got[2][ok], ref[2][ok], np.maximum(ref[2][ok], med), what='P1')
    assert_bins(got[3][ok], ref[3][ok], np.maximum(ref[3][ok], med), what='P2')
    assert_bins(got[4][ok], ref[4][ok], np.sqrt(ref[2][ok] * ref[3][ok]), what='PX')


@pytest.mark.parametrize('dims,mas', [(16, 'CIC'), (33, 'TSC'), (64, 'PCS'), (32, 'None')])
def test_correct_mas_vs_oracle(oracle, dims, mas):
    print('case correct_mas', dims, mas, flush=True)
    import torch
    from pylians3_amd import Pk_library as PKL
    rng = np.random.default_rng(dims)
    x = rng.standard_normal((dims,) * 3).astype(np.float32)
    got = PKL.correct_MAS(x, 100.0, MAS=mas)
    ref = oracle.correct_mas(x, 100.0, mas)
    ref64 = oracle.correct_mas(x.astype(np.float64), 100.0, mas, f64=True)
    rms = np.sqrt(np.mean(ref64 ** 2))
    tol = 1e-5 * np.maximum(np.abs(ref64), rms)
    assert np.all(np.abs(got - ref64) <= tol)
    assert np.all(np.abs(ref - ref64) <= tol)
    dev = PKL.correct_MAS(torch.from_numpy(x).cuda(), 100.0, MAS=mas).cpu().numpy()
    np.testing.assert_array_equal(dev, got)
