"""Drop-in Pk_library (density-field power spectra) on MI355X (gfx950).

Same names, arguments, printed messages, result attributes and error behaviour
as the reference module (library/Pk_library/Pk_library.pyx):

    frequencies, frequencies_2D        Pk_library.pyx:56-69
    MAS_function, MAS_correction        Pk_library.pyx:72-84
    check_number_modes[_2D]             Pk_library.pyx:87-114
    FFT3Dr_f(a, threads)                Pk_library.pyx:117-130   (rocFFT)
    Pk(delta, BoxSize, axis=2, MAS='CIC', threads=1, verbose=True)   :263-420
        k1D Pk1D Nmodes1D | kpar kper Pk2D Nmodes2D | k3D Pk[:,0:3] Nmodes3D Pkphase
    XPk(delta, BoxSize, axis=2, MAS=None, threads=1)                  :529-793
        k1D Pk1D PkX1D Nmodes1D | kpar kper Pk2D PkX2D Nmodes2D | k3D Pk XPk Nmodes3D
    Pk_plane(delta, BoxSize, MAS='CIC', threads=1, verbose=True)       :435-511
        k Pk Nmodes (2D maps)
    XPk_plane(delta1, delta2, BoxSize, MAS1=None, MAS2=None, threads=1) :1093-1220
        k Pk (nb,2) XPk Nmodes r
    XPk_2D(delta1, delta2, BoxSize, axis=2, MAS1='CIC', MAS2='CIC', threads=1)  :1754-1865
        [kpar, kper, Pk1, Pk2, PkX, Nmodes]
    correct_MAS(delta, BoxSize, MAS='CIC', threads=1)                  :1882-1939

``threads`` is accepted and ignored (it only fed FFTW).  Fields may be numpy
arrays or contiguous float32 torch tensors on the GPU.  The FFT, the mode loop
and the bin sums run in libpylians_hip.so; the O(bins) normalisation runs in
its host code.
"""
import ctypes
import sys
import time

import numpy as np

from . import _native as N


def frequencies(BoxSize, dims):
    kF = 2.0 * np.pi / BoxSize
    middle = dims // 2
    kN = middle * kF
    kmax_par = middle
    kmax_per = int(np.sqrt(middle ** 2 + middle ** 2))
    kmax = int(np.sqrt(middle ** 2 + middle ** 2 + middle ** 2))
    return kF, kN, kmax_par, kmax_per, kmax


def frequencies_2D(BoxSize, dims):
    kF = 2.0 * np.pi / BoxSize
    middle = dims // 2
    kN = middle * kF
    kmax_par = middle
    kmax_per = middle
    kmax = int(np.sqrt(middle ** 2 + middle ** 2))
    return kF, kN, kmax_par, kmax_per, kmax


def MAS_function(MAS):
    MAS_index = 0
    if MAS == 'NGP':
        MAS_index = 1
    if MAS == 'CIC':
        MAS_index = 2
    if MAS == 'TSC':
        MAS_index = 3
    if MAS == 'PCS':
        MAS_index = 4
    return MAS_index


def MAS_correction(x, MAS_index):
    return N.mas_correction(x, MAS_index)


def _check_modes(Nmodes, dims, own_odd, own_even, power):
    own_modes = own_odd if dims % 2 == 1 else own_even
    repeated_modes = (dims ** power - own_modes) // 2
    indep_modes = repeated_modes + own_modes
    if int(np.sum(Nmodes)) != indep_modes:
        print('WARNING: Not all modes counted')
        print('Counted  %d independent modes' % (int(np.sum(Nmodes))))
        print('Expected %d independent modes' % indep_modes)
        sys.exit()


def check_number_modes(Nmodes, dims):
    _check_modes(Nmodes, dims, 1, 8, 3)


def check_number_modes_2D(Nmodes, dims):
    _check_modes(Nmodes, dims, 1, 4, 2)


def _check_field(a, ndim=3):
    if N.is_device_array(a):
        import torch
        if a.dtype != torch.float32:
            raise ValueError("Buffer dtype mismatch, expected 'float32_t' but got '%s'" % a.dtype)
        if a.dim() != ndim:
            raise ValueError('Buffer has wrong number of dimensions (expected %d, got %d)'
                             % (ndim, a.dim()))
        if not a.is_contiguous():
            raise ValueError('device fields must be contiguous')
        return a, True
    a = np.asarray(a)
    if a.dtype != np.float32:
        raise ValueError("Buffer dtype mismatch, expected 'float32_t' but got '%s'"
                         % ('double' if a.dtype == np.float64 else a.dtype.name))
    if a.ndim != ndim:
        raise ValueError('Buffer has wrong number of dimensions (expected %d, got %d)'
                         % (ndim, a.ndim))
    return np.ascontiguousarray(a), False


def FFT3Dr_f(a, threads=1):
    """Forward unnormalised r2c FFT of a float32 cube -> complex64 (d,d,d//2+1)."""
    a, dev = _check_field(a)
    dims = len(a)
    ctx = N.default_context()
    if dev:
        import torch
        out = torch.empty((dims, dims, dims // 2 + 1), dtype=torch.complex64, device=a.device)
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_fft_r2c(ctx.handle, N.device_ptr(a), out.data_ptr(), dims,
                                        N.PTR_DEVICE))
        return out
    out = np.empty((dims, dims, dims // 2 + 1), dtype=np.complex64)
    N.check(N.lib().pyl_fft_r2c(ctx.handle, a.ctypes.data, out.ctypes.data, dims, 0))
    return out


def _run_pk(fields, mas_ids, BoxSize, axis, ctx=None, fourier=False, flags=0):
    """Call pyl_pk / pyl_pk_from_fourier; return the result arrays (spectra last)."""
    ctx = ctx or N.default_context()
    F = len(fields)
    dims = fields[0].shape[0]
    out, arrs, S = N.alloc_pk_out(dims, F)
    devs = [N.is_device_array(f) for f in fields]
    if any(devs) and not all(devs):
        raise ValueError('all fields must be on the GPU or all on the host')
    dev = all(devs)
    ptrs = (ctypes.c_void_p * F)(*[(N.device_ptr(f) if dev else f.ctypes.data) for f in fields])
    mas = (ctypes.c_int * F)(*mas_ids)
    fl = flags | (N.PTR_DEVICE if dev else 0)
    with N.torch_stream(ctx, dev):
        if fourier:
            N.check(N.lib().pyl_pk_from_fourier(ctx.handle, ptrs, F, mas, int(dims), float(BoxSize),
                                                int(axis), fl, ctypes.byref(out)))
        else:
            N.check(N.lib().pyl_pk(ctx.handle, ptrs, F, mas, int(dims), float(BoxSize), int(axis),
                                   1.0, 0.0, fl, ctypes.byref(out)))
    return arrs


class Pk:
    """1D, 2D and 3D (l = 0, 2, 4) power spectrum of a density field (Pk_library.pyx:263-420)."""

    def __init__(self, delta, BoxSize, axis=2, MAS='CIC', threads=1, verbose=True):
        start = time.time()
        axis = int(axis)
        if verbose:
            print('\nComputing power spectrum of the field...')
        delta, _ = _check_field(delta)
        dims = len(delta)
        MAS_index = MAS_function(MAS)
        start2 = time.time()
        r = _run_pk([delta], [MAS_index], BoxSize, axis)
        if verbose:
            print('Time to complete loop = %.2f' % (time.time() - start2))
        self.k1D = r['k1D']
        self.Pk1D = r['Pk1D'][:, 0].copy()
        self.Nmodes1D = r['Nmodes1D']
        self.kpar = r['kpar']
        self.kper = r['kper']
        self.Pk2D = r['Pk2D'][:, 0].copy()
        self.Nmodes2D = r['Nmodes2D']
        check_number_modes(np.concatenate([[1.0], r['Nmodes3D']]), dims)
        self.k3D = r['k3D']
        self.Nmodes3D = r['Nmodes3D']
        self.Pk = r['Pk3D'][:, :, 0].copy()
        self.Pkphase = r['Pkphase']
        if verbose:
            print('Time taken = %.2f seconds' % (time.time() - start))


class XPk:
    """Auto- and cross-power spectra of several fields (Pk_library.pyx:529-793)."""

    def __init__(self, delta, BoxSize, axis=2, MAS=None, threads=1):
        start = time.time()
        axis = int(axis)
        print('\nComputing power spectra of the fields...')
        dims = len(delta[0])
        fields = len(delta)
        for i in range(1, fields):
            if len(delta[i]) != dims:
                print('Fields have different grid sizes!!!')
                sys.exit()
        list(set(MAS))   # the reference builds its unique-MAS table with set(MAS)
        mas_ids = [MAS_function(MAS[i]) for i in range(fields)]
        checked = [_check_field(d)[0] for d in delta]
        print('Time FFTS = %.2f' % (time.time() - start))
        start2 = time.time()
        r = _run_pk(checked, mas_ids, BoxSize, axis)
        print('Time loop = %.2f' % (time.time() - start2))
        X = fields * (fields - 1) // 2
        self.k1D = r['k1D']
        self.Nmodes1D = r['Nmodes1D']
        self.Pk1D = r['Pk1D'][:, :fields].copy()
        self.PkX1D = r['Pk1D'][:, fields:fields + X].copy()
        self.kpar = r['kpar']
        self.kper = r['kper']
        self.Nmodes2D = r['Nmodes2D']
        self.Pk2D = r['Pk2D'][:, :fields].copy()
        self.PkX2D = r['Pk2D'][:, fields:fields + X].copy()
        check_number_modes(np.concatenate([[1.0], r['Nmodes3D']]), dims)
        self.k3D = r['k3D']
        self.Nmodes3D = r['Nmodes3D']
        self.Pk = r['Pk3D'][:, :, :fields].copy()
        self.XPk = r['Pk3D'][:, :, fields:fields + X].copy()
        print('Time taken = %.2f seconds' % (time.time() - start))


def Pk_from_fourier(delta_k_list, BoxSize, dims, axis=2, MAS=None):
    """Binning pass alone on half-complex fields (complex64, (d,d,d//2+1)).

    Returns a dict of raw result arrays with spectra along the last axis
    (autos first, then the cross spectra of pairs i<j).  Used to check the
    binning kernel on exactly the Fourier data another FFT produced."""
    fields = []
    for d in delta_k_list:
        if N.is_device_array(d):
            fields.append(d)
        else:
            d = np.ascontiguousarray(d, dtype=np.complex64)
            if d.shape != (dims, dims, dims // 2 + 1):
                raise ValueError('delta_k must be (dims, dims, dims//2+1)')
            fields.append(d)
    if MAS is None:
        MAS = ['None'] * len(fields)
    r = _run_pk(fields, [MAS_function(m) for m in MAS], BoxSize, axis, fourier=True)
    return r


def _plane(fields, mas_ids, BoxSize):
    """Call pyl_pk_plane on one or two (grid, grid) maps; returns (k, Nmodes, Pk (nb, S))."""
    ctx = N.default_context()
    F = len(fields)
    grid = fields[0].shape[0]
    nb = N.lib().pyl_pk_plane_nbins(int(grid))
    S = 1 if F == 1 else 3
    k = np.zeros(nb)
    Nm = np.zeros(nb)
    P = np.zeros((nb, S))
    devs = [N.is_device_array(f) for f in fields]
    if any(devs) and not all(devs):
        raise ValueError('all fields must be on the GPU or all on the host')
    dev = all(devs)
    ptrs = (ctypes.c_void_p * F)(*[(N.device_ptr(f) if dev else f.ctypes.data) for f in fields])
    mas = (ctypes.c_int * F)(*mas_ids)
    dp = ctypes.POINTER(ctypes.c_double)
    with N.torch_stream(ctx, dev):
        N.check(N.lib().pyl_pk_plane(ctx.handle, ptrs, F, mas, int(grid), float(BoxSize),
                                     N.PTR_DEVICE if dev else 0, k.ctypes.data_as(dp),
                                     Nm.ctypes.data_as(dp), P.ctypes.data_as(dp)))
    return k, Nm, P


class Pk_plane:
    """Power spectrum of a 2D map (Pk_library.pyx:435-511): attributes k, Pk, Nmodes."""

    def __init__(self, delta, BoxSize, MAS='CIC', threads=1, verbose=True):
        start = time.time()
        if verbose:
            print('\nComputing power spectrum of the field...')
        delta, _ = _check_field(delta, 2)
        MAS_index = MAS_function(MAS)
        start2 = time.time()
        try:
            k, Nm, P = _plane([delta], [MAS_index], BoxSize)
        except N.NativeError as e:
            if e.status == N.PYL_EMODES:   # check_number_modes_2D (Pk_library.pyx:102-114)
                print(str(e))
                sys.exit()
            raise
        if verbose:
            print('Time to complete loop = %.2f' % (time.time() - start2))
        self.k = k
        self.Nmodes = Nm
        self.Pk = P[:, 0].copy()
        if verbose:
            print('Time taken = %.2f seconds' % (time.time() - start))


class XPk_plane:
    """Auto- and cross-power spectra of two 2D maps (Pk_library.pyx:1093-1220):
    attributes k, Pk (nb, 2), XPk, Nmodes, r."""

    def __init__(self, delta1, delta2, BoxSize, MAS1=None, MAS2=None, threads=1):
        start = time.time()
        print('\nComputing power spectra of the fields...')
        if delta1.shape[0] != delta2.shape[1]:
            raise Exception('Images have different grid sizes!!!')
        d1, _ = _check_field(delta1, 2)
        d2, _ = _check_field(delta2, 2)
        print('Time FFTS = %.2f' % (time.time() - start))
        start2 = time.time()
        k, Nm, P = _plane([d1, d2], [MAS_function(MAS1), MAS_function(MAS2)], BoxSize)
        print('Time loop = %.2f' % (time.time() - start2))
        self.k = k
        self.Nmodes = Nm
        self.Pk = P[:, 0:2].copy()
        self.XPk = P[:, 2].copy()
        self.r = self.XPk / np.sqrt(self.Pk[:, 0] * self.Pk[:, 1])
        print('Time taken = %.2f seconds' % (time.time() - start))


def XPk_2D(delta1, delta2, BoxSize, axis=2, MAS1='CIC', MAS2='CIC', threads=1):
    """2D (k_par, k_per) auto and cross spectra of two fields (Pk_library.pyx:1754-1865).
    Returns [kpar, kper, Pk1, Pk2, PkX, Nmodes]: the 2D bins of the 3D mode loop with
    the DC bin kept, Pk*(L/d^2)^3/Nmodes."""
    start = time.time()
    print('Computing power spectra of the fields...')
    dims = len(delta1)
    if dims != len(delta2):
        print('Different grids in the two fields!!!')
        sys.exit()
    d1, _ = _check_field(delta1)
    d2, _ = _check_field(delta2)
    start2 = time.time()
    r = _run_pk([d1, d2], [MAS_function(MAS1), MAS_function(MAS2)], BoxSize, int(axis))
    print('Time compute modulus = %.2f' % (time.time() - start2))
    Pk2D, PkX2D = r['Pk2D'], r['Pk2D'][:, 2]
    print('Time taken = %.2f seconds' % (time.time() - start))
    return [r['kpar'], r['kper'], Pk2D[:, 0].copy(), Pk2D[:, 1].copy(), PkX2D.copy(),
            r['Nmodes2D']]


def correct_MAS(delta, BoxSize, MAS='CIC', threads=1):
    """Deconvolve the MAS window from a field (Pk_library.pyx:1882-1939).  Like the
    reference (FFTW_BACKWARD, unnormalised) the result is dims^3 times the
    deconvolved field, float32 (dims, dims, dims)."""
    start = time.time()
    print('\nComputing power spectrum of the field...')
    delta, dev = _check_field(delta)
    dims = len(delta)
    MAS_index = MAS_function(MAS)
    ctx = N.default_context()
    if dev:
        import torch
        out = torch.empty_like(delta)
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_correct_mas(ctx.handle, N.device_ptr(delta), N.device_ptr(out), dims,
                                            float(BoxSize), MAS_index, N.PTR_DEVICE))
    else:
        out = np.empty((dims, dims, dims), dtype=np.float32)
        N.check(N.lib().pyl_correct_mas(ctx.handle, delta.ctypes.data, out.ctypes.data, dims,
                                        float(BoxSize), MAS_index, 0))
    print('Time taken = %.2f seconds' % (time.time() - start))
    return out
