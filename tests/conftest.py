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


def assert_cells(g, g_ref, rtol=CELL_RTOL, what='grid'):
    g = np.asarray(g, dtype=np.float64)
    g_ref = np.asarray(g_ref, dtype=np.float64)
    assert g.shape == g_ref.shape, (g.shape, g_ref.shape)
    floor = np.mean(np.abs(g_ref)) if g_ref.size else 0.0
    tol = rtol * np.maximum(np.abs(g_ref), floor)
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
