"""ORACLE -- test infrastructure only.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module, and only as the checker.  The product path
(``pylians3_amd``) never imports it.

CPU restatement of the reference density-field -> power-spectrum path:

* ``ma_f32``   MAS_library.MA semantics (library/MAS_library/MAS_library.pyx:57-112,
               123-545), serial float32, via ``ma_oracle.c``.
* ``ma_f64``   the same footprints in float64 (the FP64 oracle).
* ``pk``       Pk_library.Pk (library/Pk_library/Pk_library.pyx:263-420): FFT, the
               mode loop in ``pk_oracle.c``, and the normalisation below.
* ``xpk``      Pk_library.XPk (Pk_library.pyx:529-793).
* ``ma_reference`` the reference's own MAS_c.c (MAS_c.c:8-266) compiled from
               /root/reference into oracle/_ref/ by ``make ref``.

FFT.  The reference calls pyFFTW/FFTW3 in single precision (Pk_library.pyx:117-130,
version unpinned: setup.py:137).  pyFFTW is not installed, so the float32 oracle
uses ``scipy.fft.rfftn`` on float32 input (pocketfft, single precision, forward,
unnormalised, e^{-i}); the FP64 oracle uses it in double.  The FFT is pinned in
tests/test_oracle_pk.py by a direct DFT on tiny grids, plane waves and Parseval.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = {}

MAS_INDEX = {'NGP': 1, 'CIC': 2, 'TSC': 3, 'PCS': 4}


def build(ref=True):
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    subprocess.check_call(['make', '-s', '-C', HERE])
    if ref and os.path.isdir('/root/reference/library/MAS_library'):
        subprocess.check_call(['make', '-s', '-C', HERE, 'ref'])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, 'liboracle.so')
        if not os.path.exists(path):
            build(ref=False)
        L = ctypes.CDLL(path)
        fp, dp, ip = (ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double),
                      ctypes.POINTER(ctypes.c_int))
        L.ora_ma_f32.argtypes = [fp, ctypes.c_long, ctypes.c_int, fp, fp, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.ora_ma_f64.argtypes = [fp, ctypes.c_long, ctypes.c_int, fp, dp, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int]
        L.ora_ma_f64_threads.argtypes = [fp, ctypes.c_long, ctypes.c_int, fp, dp, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.ora_axis_weights.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                       ctypes.c_int, ip, fp]
        L.ora_axis_weights.restype = ctypes.c_int
        L.ora_pk_loop.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ip] + [dp] * 9
        L.ora_pk_loop_threads.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ip] + [dp] * 9 + [ctypes.c_int]
        L.ora_mas_correction.argtypes = [ctypes.c_double, ctypes.c_int]
        L.ora_mas_correction.restype = ctypes.c_double
        L.ora_pos_redshift_space.argtypes = [fp, fp, ctypes.c_long, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_float, ctypes.c_int]
        L.ora_cic_interp.argtypes = [fp, ctypes.c_int, ctypes.c_float, fp, ctypes.c_long, fp]
        L.ora_pk_plane_loop.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ip, dp, dp, dp]
        L.ora_correct_mas_modes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


# ----------------------------------------------------------------------------- MA
def ma_f32(pos, number, BoxSize, MAS='CIC', W=None, cython2d=True, renormalize_2D=True):
    """In-place float32 MA with the reference's serial semantics (+= into number).

    2D follows MA()'s expand_dims path (MAS_library.pyx:84-110): each deposit is
    repeated F times and the whole grid is then divided by F when renormalize_2D.
    """
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    assert number.dtype == np.float32 and number.flags['C_CONTIGUOUS']
    coord = pos.shape[1]
    dims = number.shape[0]
    Wc = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    lib().ora_ma_f32(_p(pos, ctypes.c_float), pos.shape[0], coord, _p(Wc, ctypes.c_float),
                     _p(number, ctypes.c_float), dims, float(BoxSize), MAS_INDEX[MAS],
                     1 if cython2d else 0)
    if coord == 2 and cython2d and renormalize_2D and MAS != 'NGP':
        number /= float(MAS_INDEX[MAS])   # numpy float32 in-place division, as the reference
    return number


def ma_f64(pos, dims, BoxSize, MAS='CIC', W=None, threads=1):
    """Float64 oracle grid (positions/weights widened from float32).  threads > 1
    adds with OpenMP atomics (order varies at the 1e-16 level)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    coord = pos.shape[1]
    number = np.zeros((dims,) * coord, dtype=np.float64)
    Wc = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    args = (_p(pos, ctypes.c_float), pos.shape[0], coord, _p(Wc, ctypes.c_float),
            _p(number, ctypes.c_double), dims, float(BoxSize), MAS_INDEX[MAS])
    if threads > 1:
        lib().ora_ma_f64_threads(*args, int(threads))
    else:
        lib().ora_ma_f64(*args)
    return number


def axis_weights(x, inv_cell_size, dims, MAS):
    idx = (ctypes.c_int * 4)()
    C = (ctypes.c_float * 4)()
    F = lib().ora_axis_weights(ctypes.c_float(x), ctypes.c_float(inv_cell_size), dims,
                               MAS_INDEX[MAS], idx, C)
    return list(idx)[:F], np.array(list(C)[:F], dtype=np.float32)


def ref_available(flavour='strict'):
    return os.path.exists(os.path.join(HERE, '_ref', 'libmas_c_%s.so' % flavour))


def ma_reference(pos, number, BoxSize, MAS='CIC', W=None, flavour='strict', threads=1):
    """The reference's own MAS_c.c kernel (MAS_c.h:3-10), compiled into oracle/_ref."""
    if flavour not in _REF:
        L = ctypes.CDLL(os.path.join(HERE, '_ref', 'libmas_c_%s.so' % flavour))
        fp = ctypes.POINTER(ctypes.c_float)
        for name in ('NGP', 'CIC', 'TSC', 'PCS'):
            f = getattr(L, name)
            f.argtypes = [fp, fp, fp, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                          ctypes.c_float, ctypes.c_int]
            f.restype = None
        _REF[flavour] = L
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    Wc = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    getattr(_REF[flavour], MAS)(_p(pos, ctypes.c_float), _p(number, ctypes.c_float),
                                _p(Wc, ctypes.c_float), pos.shape[0], number.shape[0],
                                pos.shape[1], float(BoxSize), int(threads))
    return number


_CY = {}


def cython_available(flavour='strict'):
    import glob
    return bool(glob.glob(os.path.join(HERE, '_ref', 'cy_%s' % flavour, 'MAS_library*.so')))


def cython_module(flavour='strict'):
    """The reference's own Cython MAS_library (MAS_library.pyx + MAS_c.c), translated
    and compiled from the reference sources into oracle/_ref by `make cython`."""
    if flavour not in _CY:
        import glob
        import importlib.util
        path = glob.glob(os.path.join(HERE, '_ref', 'cy_%s' % flavour, 'MAS_library*.so'))[0]
        spec = importlib.util.spec_from_file_location('MAS_library', path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        _CY[flavour] = mod
    return _CY[flavour]


def ma_cython(pos, number, BoxSize, MAS='CIC', W=None, renormalize_2D=True, flavour='strict'):
    """MA() of the reference module itself (MAS_library.pyx:57-112): serial loops."""
    cython_module(flavour).MA(pos, number, BoxSize, MAS=MAS, W=W, verbose=False,
                              renormalize_2D=renormalize_2D)
    return number


def cic_interp(density, BoxSize, pos):
    """MAS_library.CIC_interp (MAS_library.pyx:558-599): float32 gather."""
    density = np.ascontiguousarray(density, dtype=np.float32)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    den = np.zeros(pos.shape[0], np.float32)
    lib().ora_cic_interp(_p(density, ctypes.c_float), density.shape[0], float(BoxSize),
                         _p(pos, ctypes.c_float), pos.shape[0], _p(den, ctypes.c_float))
    return den


def pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    vel = np.ascontiguousarray(vel, dtype=np.float32)
    lib().ora_pos_redshift_space(_p(pos, ctypes.c_float), _p(vel, ctypes.c_float),
                                 pos.shape[0], BoxSize, Hubble, redshift, axis)
    return pos


# ----------------------------------------------------------------------------- Pk
def frequencies(BoxSize, dims):
    """Pk_library.pyx:56-61."""
    kF = 2.0 * np.pi / BoxSize
    middle = dims // 2
    kN = middle * kF
    kmax_par = middle
    kmax_per = int(np.sqrt(middle ** 2 + middle ** 2))
    kmax = int(np.sqrt(middle ** 2 + middle ** 2 + middle ** 2))
    return kF, kN, kmax_par, kmax_per, kmax


def mas_correction(x, p):
    return lib().ora_mas_correction(float(x), int(p))


def fft_f32(delta, workers=None):
    import scipy.fft
    return scipy.fft.rfftn(np.asarray(delta, dtype=np.float32), workers=workers)


def fft_f64(delta, workers=None):
    import scipy.fft
    return scipy.fft.rfftn(np.asarray(delta, dtype=np.float64), workers=workers)


def expected_modes(dims):
    """check_number_modes closed form (Pk_library.pyx:87-99)."""
    own = 1 if dims % 2 == 1 else 8
    return (dims ** 3 - own) // 2 + own


def raw_bins(dk_list, dims, axis, mas_list, is_f64=False, threads=1):
    """Run the reference mode loop on F half-complex fields; return raw sums.
    threads > 1 splits the loop over kx into threads with private sums added in
    thread order (deterministic; sums differ from the serial loop at 1e-16)."""
    F = len(dk_list)
    S = F + F * (F - 1) // 2
    middle = dims // 2
    kmax_par = middle
    kmax_per = int(np.sqrt(2 * middle ** 2))
    kmax = int(np.sqrt(3 * middle ** 2))
    nb2 = (kmax_par + 1) * (kmax_per + 1)
    ct = np.complex128 if is_f64 else np.complex64
    if F == 1:      # no stacked copy (a 1024^3 complex128 field is 8.6 GB)
        dk = np.ascontiguousarray(np.asarray(dk_list[0], dtype=ct))[None]
    else:
        dk = np.ascontiguousarray(np.stack([np.asarray(d, dtype=ct) for d in dk_list]))
    assert dk.shape[1:] == (dims, dims, middle + 1)
    out = dict(
        k1D=np.zeros(kmax_par + 1), Pk1D=np.zeros((kmax_par + 1, S)), N1D=np.zeros(kmax_par + 1),
        Pk2D=np.zeros((nb2, S)), N2D=np.zeros(nb2),
        k3D=np.zeros(kmax + 1), Pk3D=np.zeros((kmax + 1, 3, S)), Pkphase=np.zeros(kmax + 1),
        N3D=np.zeros(kmax + 1))
    mas = (ctypes.c_int * F)(*[MAS_INDEX.get(m, 0) if isinstance(m, str) else int(m)
                               for m in mas_list])
    args = (dk.ctypes.data, 1 if is_f64 else 0, F, dims, axis, mas,
            _p(out['k1D'], ctypes.c_double), _p(out['Pk1D'], ctypes.c_double),
            _p(out['N1D'], ctypes.c_double), _p(out['Pk2D'], ctypes.c_double),
            _p(out['N2D'], ctypes.c_double), _p(out['k3D'], ctypes.c_double),
            _p(out['Pk3D'], ctypes.c_double), _p(out['Pkphase'], ctypes.c_double),
            _p(out['N3D'], ctypes.c_double))
    if threads > 1:
        lib().ora_pk_loop_threads(*args, int(threads))
    else:
        lib().ora_pk_loop(*args)
    return out


class _Result:
    pass


def finalize_pk(raw, dims, BoxSize, spectra):
    """Normalisation of Pk_library.pyx:384-418 / :735-791 for the given spectra columns."""
    kF, kN, kmax_par, kmax_per, kmax = frequencies(BoxSize, dims)
    fact = (BoxSize / dims ** 2) ** 3
    r = _Result()
    # 1D (Pk_library.pyx:384-391): drop DC bin
    k1D = raw['k1D'][1:].copy()
    N1D = raw['N1D'][1:].copy()
    Pk1D = raw['Pk1D'][1:, spectra].copy()
    for i in range(len(k1D)):
        Pk1D[i] = Pk1D[i] * fact
        k1D[i] = (k1D[i] / N1D[i]) * kF
        kmaxper = np.sqrt(kN ** 2 - k1D[i] ** 2)
        Pk1D[i] = Pk1D[i] * (np.pi * kmaxper ** 2 / N1D[i]) / (2.0 * np.pi) ** 2
    r.k1D, r.Pk1D, r.Nmodes1D = k1D, Pk1D, N1D
    # 2D (Pk_library.pyx:394-404): keep all bins, bin centres
    nb2 = (kmax_par + 1) * (kmax_per + 1)
    kpar = np.zeros(nb2)
    kper = np.zeros(nb2)
    for k_par in range(kmax_par + 1):
        for k_per in range(kmax_per + 1):
            index_2D = (kmax_par + 1) * k_per + k_par
            kpar[index_2D] = 0.5 * (k_par + k_par + 1) * kF
            kper[index_2D] = 0.5 * (k_per + k_per + 1) * kF
    N2D = raw['N2D'].copy()
    Pk2D = raw['Pk2D'][:, spectra] * fact / N2D[:, None]
    r.kpar, r.kper, r.Pk2D, r.Nmodes2D = kpar, kper, Pk2D, N2D
    # 3D (Pk_library.pyx:408-418)
    assert int(np.sum(raw['N3D'])) == expected_modes(dims)
    N3D = raw['N3D'][1:].copy()
    k3D = (raw['k3D'][1:] / N3D) * kF
    P = raw['Pk3D'][1:, :, spectra].copy()
    P[:, 0] = (P[:, 0] / N3D[:, None]) * fact
    P[:, 1] = (P[:, 1] * 5.0 / N3D[:, None]) * fact
    P[:, 2] = (P[:, 2] * 9.0 / N3D[:, None]) * fact
    r.k3D, r.Nmodes3D, r.Pk = k3D, N3D, P
    r.Pkphase = (raw['Pkphase'][1:] / N3D) * fact
    return r


def pk(delta, BoxSize, axis=2, MAS='CIC', f64=False, dk=None, workers=None, threads=1):
    """Pk_library.Pk restated (float32 FFT + reference loop) or the FP64 oracle.

    Returns an object with the reference attributes (k3D, Pk (kmax,3), Nmodes3D,
    k1D, Pk1D, Nmodes1D, kpar, kper, Pk2D, Nmodes2D, Pkphase).
    """
    dims = delta.shape[0]
    if dk is None:
        dk = fft_f64(delta, workers) if f64 else fft_f32(delta, workers)
    raw = raw_bins([dk], dims, axis, [MAS], is_f64=f64, threads=threads)
    r = finalize_pk(raw, dims, BoxSize, [0])
    r.Pk = r.Pk[:, :, 0]
    r.Pk1D = r.Pk1D[:, 0]
    r.Pk2D = r.Pk2D[:, 0]
    r.raw = raw
    return r


def xpk(deltas, BoxSize, axis=2, MAS=None, f64=False, dks=None, workers=None, threads=1):
    """Pk_library.XPk restated.  Pk (kmax,3,F), XPk (kmax,3,F(F-1)/2), etc."""
    F = len(deltas)
    dims = deltas[0].shape[0]
    if dks is None:
        dks = [fft_f64(d, workers) if f64 else fft_f32(d, workers) for d in deltas]
    raw = raw_bins(dks, dims, axis, MAS, is_f64=f64, threads=threads)
    autos = list(range(F))
    crosses = list(range(F, F + F * (F - 1) // 2))
    a = finalize_pk(raw, dims, BoxSize, autos)
    r = _Result()
    r.k3D, r.Nmodes3D, r.k1D, r.Nmodes1D = a.k3D, a.Nmodes3D, a.k1D, a.Nmodes1D
    r.kpar, r.kper, r.Nmodes2D = a.kpar, a.kper, a.Nmodes2D
    r.Pk, r.Pk1D, r.Pk2D = a.Pk, a.Pk1D, a.Pk2D
    if crosses:
        c = finalize_pk(raw, dims, BoxSize, crosses)
        r.XPk, r.PkX1D, r.PkX2D = c.Pk, c.Pk1D, c.Pk2D
    else:
        r.XPk = np.zeros((len(a.k3D), 3, 0))
        r.PkX1D = np.zeros((len(a.k1D), 0))
        r.PkX2D = np.zeros((len(a.kpar), 0))
    r.raw = raw
    return r


# ----------------------------------------------------------------------------- brute force
def brute_modes(dims, axis):
    """Enumerate all dims^3 modes and keep one of each Hermitian pair (plus the
    self-conjugate ones): the independent modes, found without the reference
    loop's half-space skip rule.  Returns a dict of per-mode integer arrays
    kx, ky, kz (signed; a component equal to middle stays +middle for even dims),
    k2, kidx = floor(|k|), kpar = |k_par|, kper = floor(|k_perp|), i2 (2D bin)."""
    middle = dims // 2
    kmax_par = middle
    k = np.arange(dims)
    k = np.where(k > middle, k - dims, k)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing='ij')
    # canonical representative of each pair: compare the flat index of v and -v (mod dims)
    fi = (KX % dims) * dims * dims + (KY % dims) * dims + (KZ % dims)
    fj = ((-KX) % dims) * dims * dims + ((-KY) % dims) * dims + ((-KZ) % dims)
    keep = fi <= fj
    kx, ky, kz = KX[keep], KY[keep], KZ[keep]
    k2 = kx * kx + ky * ky + kz * kz
    kidx = np.floor(np.sqrt(k2)).astype(np.int64)
    par = [kx, ky, kz][axis]
    perp2 = k2 - par * par
    kper = np.floor(np.sqrt(perp2)).astype(np.int64)
    kpar = np.abs(par)
    return dict(kx=kx, ky=ky, kz=kz, k2=k2, kidx=kidx, kpar=kpar, kper=kper,
                i2=(kmax_par + 1) * kper + kpar)


def brute_mode_counts(dims, axis):
    """Histogram of the independent modes (brute_modes) into the reference's
    1D/2D/3D bins: (N1D, N2D, N3D)."""
    middle = dims // 2
    kmax_par, kmax_per = middle, int(np.sqrt(2 * middle ** 2))
    kmax = int(np.sqrt(3 * middle ** 2))
    m = brute_modes(dims, axis)
    N3D = np.bincount(m['kidx'], minlength=kmax + 1).astype(np.float64)
    N2D = np.bincount(m['i2'], minlength=(kmax_par + 1) * (kmax_per + 1)).astype(np.float64)
    m1 = m['k2'] <= middle * middle
    N1D = np.bincount(m['kpar'][m1], minlength=kmax_par + 1).astype(np.float64)
    return N1D, N2D, N3D


def brute_k3d_sums(dims, axis=0):
    """Sum of |k| (grid units, float64 sqrt of the integer k^2) over the
    independent modes of each 3D bin."""
    kmax = int(np.sqrt(3 * (dims // 2) ** 2))
    m = brute_modes(dims, axis)
    return np.bincount(m['kidx'], weights=np.sqrt(m['k2'].astype(np.float64)), minlength=kmax + 1)


def direct_dft_r2c(x):
    """O(d^6) forward DFT of a real cube, half-complex output (tiny grids only)."""
    d = x.shape[0]
    n = np.arange(d)
    E = np.exp(-2j * np.pi * np.outer(n, n) / d)
    X = np.einsum('ai,bj,ck,ijk->abc', E, E, E, x.astype(np.complex128))
    return X[:, :, :d // 2 + 1]


# ----------------------------------------------------------------------------- 2D spectra
def frequencies_2D(BoxSize, grid):
    """Pk_library.pyx:64-69."""
    kF = 2.0 * np.pi / BoxSize
    middle = grid // 2
    kN = middle * kF
    return kF, kN, middle, middle, int(np.sqrt(middle ** 2 + middle ** 2))


def expected_modes_2D(grid):
    """check_number_modes_2D closed form (Pk_library.pyx:102-114)."""
    own = 1 if grid % 2 == 1 else 4
    return (grid ** 2 - own) // 2 + own


def fft2_f32(delta):
    import scipy.fft
    return scipy.fft.rfftn(np.asarray(delta, dtype=np.float32))


def fft2_f64(delta):
    import scipy.fft
    return scipy.fft.rfftn(np.asarray(delta, dtype=np.float64))


def raw_plane_bins(dk_list, grid, mas_list, is_f64=False):
    """Mode loop of Pk_plane (one field) or XPk_plane (two fields): raw sums."""
    F = len(dk_list)
    S = 1 if F == 1 else 3
    kmax = int(np.sqrt(2 * (grid // 2) ** 2))
    ct = np.complex128 if is_f64 else np.complex64
    if F == 1:      # no stacked copy (a 1024^3 complex128 field is 8.6 GB)
        dk = np.ascontiguousarray(np.asarray(dk_list[0], dtype=ct))[None]
    else:
        dk = np.ascontiguousarray(np.stack([np.asarray(d, dtype=ct) for d in dk_list]))
    assert dk.shape[1:] == (grid, grid // 2 + 1)
    out = dict(k=np.zeros(kmax + 1), Pk=np.zeros((kmax + 1, S)), N=np.zeros(kmax + 1))
    mas = (ctypes.c_int * F)(*[MAS_INDEX.get(m, 0) if isinstance(m, str) or m is None else int(m)
                               for m in mas_list])
    lib().ora_pk_plane_loop(dk.ctypes.data, 1 if is_f64 else 0, F, grid, mas,
                            _p(out['k'], ctypes.c_double), _p(out['Pk'], ctypes.c_double),
                            _p(out['N'], ctypes.c_double))
    return out


def pk_plane(delta, BoxSize, MAS='CIC', f64=False):
    """Pk_library.Pk_plane restated: attributes k, Pk, Nmodes (Pk_library.pyx:499-507)."""
    grid = delta.shape[0]
    kF = 2.0 * np.pi / BoxSize
    dk = fft2_f64(delta) if f64 else fft2_f32(delta)
    raw = raw_plane_bins([dk], grid, [MAS], f64)
    assert int(np.sum(raw['N'])) == expected_modes_2D(grid)
    r = _Result()
    N = raw['N'][1:].copy()
    r.k = (raw['k'][1:] / N) * kF
    r.Pk = (raw['Pk'][1:, 0] / N) * (BoxSize / grid ** 2) ** 2
    r.Nmodes = N
    r.raw = raw
    return r


def xpk_plane(delta1, delta2, BoxSize, MAS1=None, MAS2=None, f64=False):
    """Pk_library.XPk_plane restated: k, Pk (nb,2), XPk, Nmodes, r (Pk_library.pyx:1205-1218)."""
    grid = delta1.shape[0]
    kF = 2.0 * np.pi / BoxSize
    dks = [fft2_f64(d) if f64 else fft2_f32(d) for d in (delta1, delta2)]
    raw = raw_plane_bins(dks, grid, [MAS1, MAS2], f64)
    fact = (BoxSize / grid ** 2) ** 2
    N = raw['N'][1:].copy()
    r = _Result()
    r.k = (raw['k'][1:] / N) * kF
    r.Pk = (raw['Pk'][1:, 0:2] / N[:, None]) * fact
    r.XPk = (raw['Pk'][1:, 2] / N) * fact
    r.Nmodes = N
    r.r = r.XPk / np.sqrt(r.Pk[:, 0] * r.Pk[:, 1])
    r.raw = raw
    return r


def xpk_2d(delta1, delta2, BoxSize, axis=2, MAS1='CIC', MAS2='CIC', f64=False):
    """Pk_library.XPk_2D restated: [kpar, kper, Pk1, Pk2, PkX, Nmodes]
    (Pk_library.pyx:1754-1865; the 2D bins of the 3D mode loop, kept with the DC bin)."""
    dims = delta1.shape[0]
    dks = [fft_f64(d) if f64 else fft_f32(d) for d in (delta1, delta2)]
    raw = raw_bins(dks, dims, axis, [MAS1, MAS2], is_f64=f64)
    kF, kN, kmax_par, kmax_per, kmax = frequencies(1.0, dims)
    kF = 2.0 * np.pi / BoxSize
    nb2 = (kmax_par + 1) * (kmax_per + 1)
    kpar = np.zeros(nb2)
    kper = np.zeros(nb2)
    for ipar in range(kmax_par + 1):
        for iper in range(kmax_per + 1):
            index = (kmax_par + 1) * iper + ipar
            kpar[index] = 0.5 * (ipar + ipar + 1) * kF
            kper[index] = 0.5 * (iper + iper + 1) * kF
    N = raw['N2D'].copy()
    fact = (BoxSize / dims ** 2) ** 3
    with np.errstate(divide='ignore', invalid='ignore'):
        Pk1 = raw['Pk2D'][:, 0] * fact / N
        Pk2 = raw['Pk2D'][:, 1] * fact / N
        PkX = raw['Pk2D'][:, 2] * fact / N
    return [kpar, kper, Pk1, Pk2, PkX, N]


def correct_mas(delta, BoxSize, MAS='CIC', f64=False):
    """Pk_library.correct_MAS restated (Pk_library.pyx:1882-1936): r2c FFT, MAS
    deconvolution of the independent modes, unnormalised c2r (FFTW_BACKWARD).

    The inverse real transform reads the redundant half of the kz = 0 and
    kz = middle planes through its Hermitian part, so each of those modes ends
    up with (1 + c)/2 of the correction c; scipy's irfftn and FFTW both behave
    this way (a complex transform over x, y followed by a c2r along z that
    uses only the real part of the kz = 0 and kz = middle terms)."""
    import scipy.fft
    dims = delta.shape[0]
    dk = np.ascontiguousarray(fft_f64(delta) if f64 else fft_f32(delta))
    lib().ora_correct_mas_modes(dk.ctypes.data, 1 if f64 else 0, dims,
                                MAS_INDEX.get(MAS, 0) if isinstance(MAS, str) else int(MAS))
    out = scipy.fft.irfftn(dk, s=delta.shape, norm='forward')
    return out.astype(np.float64 if f64 else np.float32)


def brute_mode_counts_2D(grid):
    """Enumerate all grid^2 modes, keep one of each Hermitian pair, histogram by int(|k|)."""
    middle = grid // 2
    kmax = int(np.sqrt(2 * middle ** 2))
    k = np.arange(grid)
    k = np.where(k > middle, k - grid, k)
    KX, KY = np.meshgrid(k, k, indexing='ij')
    fi = (KX % grid) * grid + (KY % grid)
    fj = ((-KX) % grid) * grid + ((-KY) % grid)
    keep = fi <= fj
    k2 = KX[keep] ** 2 + KY[keep] ** 2
    return np.bincount(np.floor(np.sqrt(k2)).astype(np.int64), minlength=kmax + 1).astype(np.float64)
