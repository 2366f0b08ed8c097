import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP path)')
    config.addinivalue_line('markers', 'slow: full-size case (seconds of GPU time)')


# ---------------------------------------------------------------- tolerances
CELL_RTOL = 1e-5   # per cell: |g - g_ref| <= 1e-5 * max(|g_ref|, mean |g_ref|)
BIN_RTOL = 1e-5    # per bin:  |P - P_ref| <= 1e-5 * max(|P_ref|, floor)


U32 = 2.0 ** -24    # float32 unit roundoff


def contrib_counts(pos, dims, box, mas, cython2d=True, chunk=1 << 22):
    """Number of float32 additions the reference makes into each cell (its serial
    loop adds every footprint weight, zero or not; the 2D expand_dims path adds
    each (l, m) weight F times).  Bounds the reference's own summation error:
    |fl(sum) - sum| <= n_adds * u * sum|x_i|.  Particles are taken in chunks so
    that 10^8-particle inputs fit in memory."""
    F = {'NGP': 1, 'CIC': 2, 'TSC': 3, 'PCS': 4}[mas]
    coord = pos.shape[1]
    cnt = np.zeros(dims ** coord, np.float64)
    for s in range(0, pos.shape[0], chunk):
        x = pos[s:s + chunk].astype(np.float64) * (dims / box)
        if mas == 'NGP':
            base = np.floor(x + 0.5)
        elif mas == 'CIC':
            base = np.floor(x)
        else:
            base = np.floor(x - (F / 2.0)) + 1
        base = base.astype(np.int64)
        idx = [(base[:, ax, None] + np.arange(F)[None, :]) % dims for ax in range(coord)]  # (n, F)
        if coord == 3:
            cells = ((idx[0][:, :, None, None] * dims + idx[1][:, None, :, None]) * dims
                     + idx[2][:, None, None, :])
        else:
            cells = idx[0][:, :, None] * dims + idx[1][:, None, :]
        cnt += np.bincount(cells.ravel(), minlength=dims ** coord)
    if coord == 2 and cython2d:
        cnt *= F
    return cnt.reshape((dims,) * coord)


def assert_cells(g, g_ref, rtol=CELL_RTOL, what='grid', nadds=None):
    """Per cell: |g - g_ref| <= 1e-5 max(|g_ref|, mean|g_ref|), widened where the
    reference's serial float32 sum itself carries more error than that
    ((nadds + 2) u |g_ref| for non-negative contributions)."""
    g = np.asarray(g, dtype=np.float64)
    g_ref = np.asarray(g_ref, dtype=np.float64)
    assert g.shape == g_ref.shape, (g.shape, g_ref.shape)
    floor = np.mean(np.abs(g_ref)) if g_ref.size else 0.0
    tol = rtol * np.maximum(np.abs(g_ref), floor)
    if nadds is not None:
        tol = np.maximum(tol, (np.asarray(nadds) + 2.0) * U32 * np.abs(g_ref))
    err = np.abs(g - g_ref)
    bad = err > tol
    if bad.any():
        i = np.argmax(err - tol)
        raise AssertionError('%s: %d/%d cells out of tolerance; worst flat index %d: got %r ref %r'
                             % (what, int(bad.sum()), g.size, i, g.ravel()[i], g_ref.ravel()[i]))


def assert_bins(P, P_ref, floor, rtol=BIN_RTOL, what='Pk'):
    P = np.asarray(P, dtype=np.float64)
    P_ref = np.asarray(P_ref, dtype=np.float64)
    assert P.shape == P_ref.shape, (what, P.shape, P_ref.shape)
    tol = rtol * np.maximum(np.abs(P_ref), np.abs(floor))
    err = np.abs(P - P_ref)
    bad = ~(err <= tol)
    if bad.any():
        i = np.argmax(np.where(np.isfinite(err), err - tol, np.inf))
        raise AssertionError('%s: %d/%d bins out of tolerance; worst %d: got %r ref %r tol %r'
                             % (what, int(bad.sum()), P.size, i, P.ravel()[i], P_ref.ravel()[i],
                                np.broadcast_to(tol, P.shape).ravel()[i]))


@pytest.fixture(scope='session')
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope='session')
def gpu_ctx():
    from pylians3_amd import _native as N
    return N.default_context()


def assert_bins_fp64(P, P_ref32, P_ref64, floor=None, rtol=BIN_RTOL, what='Pk'):
    """Per bin against the FP64 oracle: |P - P64| <= max(rtol * max(|P64|, |floor|),
    3 |P32 - P64|), where P32 is the oracle's float32 result (the reference's own
    arithmetic).  For steep spectra the float32 FFT's error, which is set by the
    field's total power, dominates the low-power bins in both implementations."""
    P = np.asarray(P, dtype=np.float64)
    P32 = np.asarray(P_ref32, dtype=np.float64)
    P64 = np.asarray(P_ref64, dtype=np.float64)
    fl = np.abs(P64) if floor is None else np.maximum(np.abs(P64), np.abs(floor))
    tol = np.maximum(rtol * fl, 3.0 * np.abs(P32 - P64))
    err = np.abs(P - P64)
    bad = ~(err <= tol)
    if bad.any():
        i = np.argmax(np.where(np.isfinite(err), err - tol, np.inf))
        raise AssertionError('%s: %d/%d bins out of tolerance; worst %d: got %r f32 ref %r f64 %r tol %r'
                             % (what, int(bad.sum()), P.size, i, P.ravel()[i], P32.ravel()[i],
                                P64.ravel()[i], tol.ravel()[i]))


def err_stats(x, ref, floor=None):
    """Max and RMS of |x - ref| / max(|ref|, floor) (floor defaults to mean |ref|)."""
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    fl = np.mean(np.abs(ref)) if floor is None else np.abs(np.asarray(floor, dtype=np.float64)).ravel()
    den = np.maximum(np.abs(ref), fl)
    num = np.abs(x - ref)
    # a bin that is exactly zero in both (e.g. Pk1D at k_par = middle) has no error
    rel = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num > 0, np.inf, 0.0))
    return float(rel.max()), float(np.sqrt(np.mean(rel * rel)))


def assert_bins_same_delta(P, P32, P64, floor, rtol=BIN_RTOL, what='Pk', factor=3.0):
    """Per bin, for spectra of one float32 delta: within rtol of the float32 oracle
    (the reference's arithmetic, Pk_library.pyx:311-418, with scipy's float32 FFT in
    place of FFTW's), or, where the two float32 FFTs differ by more than that, within
    `factor` times the float32 oracle's own error against the FP64 transform of the
    same delta.  That error is taken two ways, and the looser one applies: the
    oracle's error in this bin, and its error scale over the whole spectrum.

    A float32 FFT leaves an absolute error e in every mode amplitude, set by the
    field's total power; a bin of power s = |d|^2 then carries about 2 sqrt(s) e,
    which is far above 1e-5 of s for bins of much lower power than the field's peak
    (the corner of the cube under a MAS window, 2D bins of one or two such modes).
    Which of those few bins a given float32 FFT gets nearly right is luck, so the
    per-bin error of one float32 FFT does not bound another's; the scale
    S = max_i |P32_i - P64_i| / sqrt(s_i), s_i = max(|P64_i|, |floor_i|), does,
    and the second bound is factor * S * sqrt(s_i)."""
    P = np.asarray(P, dtype=np.float64)
    P32 = np.asarray(P32, dtype=np.float64)
    P64 = np.asarray(P64, dtype=np.float64)
    fl = np.maximum(np.abs(P32), np.abs(floor))
    ok32 = np.abs(P - P32) <= rtol * fl
    s = np.sqrt(np.maximum(np.abs(P64), np.abs(floor)))
    den = np.where(s > 0, s, 1.0)
    S = float(np.max(np.where(s > 0, np.abs(P32 - P64) / den, 0.0)))
    err64 = np.abs(P - P64)
    ok64 = err64 <= np.maximum(rtol * fl, factor * np.maximum(np.abs(P32 - P64), S * s))
    bad = ~(ok32 | ok64)
    if bad.any():
        i = int(np.argmax(np.where(bad, err64 / np.maximum(s, 1e-300), -1.0)))
        raise AssertionError('%s: %d/%d bins out of tolerance; worst %d: got %r f32 ref %r f64 %r; '
                             'error scale %.3e, float32 oracle %.3e'
                             % (what, int(bad.sum()), P.size, i, P.ravel()[i], P32.ravel()[i],
                                P64.ravel()[i], err64.ravel()[i] / max(s.ravel()[i], 1e-300), S))
