"""Drop-in MAS_library on MI355X (gfx950).

Same names, arguments, in-place ``+=`` semantics and error behaviour as the
reference module (library/MAS_library/MAS_library.pyx):

    MA(pos, number, BoxSize, MAS='CIC', W=None, verbose=False, renormalize_2D=True)
        MAS_library.pyx:57-112
    NGP/CIC/TSC/PCS(pos, number, BoxSize), NGPW/CICW/TSCW/PCSW(pos, number, BoxSize, W)
        MAS_library.pyx:123-545 (3D kernels; a (d,d,1) grid with 2D positions gives the
        expand_dims behaviour of MA)
    NGPc3D ... PCSWc2D(pos, number, [W,] BoxSize, threads)
        MAS_library.pyx:1305-1389 (the C kernels, MAS_c.c:8-266: native 2D, C-contiguous)
    FLOAT_type()                                   MAS_library.pyx:11-13

Arrays may be numpy arrays (host; staged through device memory) or contiguous
torch tensors on the GPU (device-resident; no copies).  All arithmetic runs in
libpylians_hip.so; there is no CPU path.
"""
import sys
import time

import numpy as np

from . import _native as N

_MAS_IDS = {'NGP': N.NGP, 'CIC': N.CIC, 'TSC': N.TSC, 'PCS': N.PCS}


def FLOAT_type():
    return np.float32


def _check_f32(a, ndim, name):
    """Cython typed-memoryview checks: float32 dtype and dimensionality (ValueError)."""
    if N.is_device_array(a):
        import torch
        if a.dtype != torch.float32:
            raise ValueError("Buffer dtype mismatch, expected 'float32_t' but got '%s'" % a.dtype)
        if a.dim() != ndim:
            raise ValueError('Buffer has wrong number of dimensions (expected %d, got %d)'
                             % (ndim, a.dim()))
        return
    if not isinstance(a, np.ndarray):
        raise TypeError('%s must be a numpy array or a GPU torch tensor' % name)
    if a.dtype != np.float32:
        raise ValueError("Buffer dtype mismatch, expected 'float32_t' but got '%s'"
                         % ('double' if a.dtype == np.float64 else a.dtype.name))
    if a.ndim != ndim:
        raise ValueError('Buffer has wrong number of dimensions (expected %d, got %d)'
                         % (ndim, a.ndim))


def _deposit(pos, number, BoxSize, mas, W, flags, require_contig=False, ctx=None):
    """Run pyl_ma on host or device arrays.  number is updated in place."""
    axes = pos.shape[1]
    dims = number.shape[0]
    n = pos.shape[0]
    dev = N.is_device_array(pos) or N.is_device_array(number)
    if dev:
        if not (N.is_device_array(pos) and N.is_device_array(number)
                and (W is None or N.is_device_array(W))):
            raise ValueError('pos, number and W must all be on the GPU or all on the host')
        ctx = ctx or N.default_context()
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_ma(ctx.handle, N.device_ptr(pos), n, axes,
                                   None if W is None else N.device_ptr(W), N.device_ptr(number),
                                   dims, float(BoxSize), mas, flags | N.PTR_DEVICE))
        return
    if require_contig:
        for a, nm in ((pos, 'pos'), (number, 'number'), (W, 'W')):
            if a is not None and not a.flags['C_CONTIGUOUS']:
                raise ValueError('ndarray is not C-contiguous')
    ctx = ctx or N.default_context()
    p = np.ascontiguousarray(pos)
    w = None if W is None else np.ascontiguousarray(W)
    g = number if number.flags['C_CONTIGUOUS'] else np.ascontiguousarray(number)
    N.check(N.lib().pyl_ma(ctx.handle, p.ctypes.data, n, axes,
                           None if w is None else w.ctypes.data, g.ctypes.data, dims,
                           float(BoxSize), mas, flags))
    if g is not number:
        number[...] = g


def MA(pos, number, BoxSize, MAS='CIC', W=None, verbose=False, renormalize_2D=True):
    """Mass assignment of particles (optionally weighted) onto a 2D or 3D grid, += in place.

    MAS_library.pyx:57-112.  pos (N,2|3) float32, number (d,d[,d]) float32,
    W (N,) float32 or None.
    """
    coord, coord_aux = pos.shape[1], number.ndim
    if coord != coord_aux:
        print('pos have %d dimensions and the density %d!!!' % (coord, coord_aux))
        sys.exit()
    if verbose:
        if W is None:
            print('\nUsing %s mass assignment scheme' % MAS)
        else:
            print('\nUsing %s mass assignment scheme with weights' % MAS)
    start = time.time()
    if coord not in (2, 3):       # the reference dispatches only on coord 2 and 3
        return
    if MAS not in _MAS_IDS:
        print('option not valid!!!')
        sys.exit()
    _check_f32(pos, 2, 'pos')
    _check_f32(number, coord, 'number')
    if W is not None:
        _check_f32(W, 1, 'W')
    flags = 0
    if coord == 2:
        flags |= N.MA_CYTHON_2D
        if renormalize_2D:
            flags |= N.MA_RENORM_2D
    _deposit(pos, number, BoxSize, _MAS_IDS[MAS], W, flags)
    if verbose:
        print('Time taken = %.3f seconds\n' % (time.time() - start))


def _kernel3d(mas, pos, number, BoxSize, W=None):
    # The Cython kernels take a 3D grid; with 2D positions on a (d,d,1) grid the
    # unused axis keeps weight 1 and index 0, so every deposit repeats F times.
    _check_f32(pos, 2, 'pos')
    _check_f32(number, 3, 'number')
    if W is not None:
        _check_f32(W, 1, 'W')
    if pos.shape[1] == 2:
        if number.shape[2] != 1:
            raise ValueError('2D positions need a (d,d,1) grid')
        g = number.reshape(number.shape[0], number.shape[1]) if not N.is_device_array(number) \
            else number.view(number.shape[0], number.shape[1])
        _deposit(pos, g, BoxSize, mas, W, N.MA_CYTHON_2D)
        return
    _deposit(pos, number, BoxSize, mas, W, 0)


def NGP(pos, number, BoxSize):
    _kernel3d(N.NGP, pos, number, BoxSize)


def CIC(pos, number, BoxSize):
    _kernel3d(N.CIC, pos, number, BoxSize)


def TSC(pos, number, BoxSize):
    _kernel3d(N.TSC, pos, number, BoxSize)


def PCS(pos, number, BoxSize):
    _kernel3d(N.PCS, pos, number, BoxSize)


def NGPW(pos, number, BoxSize, W):
    _kernel3d(N.NGP, pos, number, BoxSize, W)


def CICW(pos, number, BoxSize, W):
    _kernel3d(N.CIC, pos, number, BoxSize, W)


def TSCW(pos, number, BoxSize, W):
    _kernel3d(N.TSC, pos, number, BoxSize, W)


def PCSW(pos, number, BoxSize, W):
    _kernel3d(N.PCS, pos, number, BoxSize, W)


def _ckernel(mas, pos, number, BoxSize, W, ndim):
    # MAS_library.pyx:1305-1389: FLOAT[:,::1] memoryviews (C-contiguous), native 2D
    _check_f32(pos, 2, 'pos')
    _check_f32(number, ndim, 'number')
    if W is not None:
        _check_f32(W, 1, 'W')
    _deposit(pos, number, BoxSize, mas, W, 0, require_contig=True)


def NGPc3D(pos, number, BoxSize, threads):
    _ckernel(N.NGP, pos, number, BoxSize, None, 3)


def NGPWc3D(pos, number, W, BoxSize, threads):
    _ckernel(N.NGP, pos, number, BoxSize, W, 3)


def NGPc2D(pos, number, BoxSize, threads):
    _ckernel(N.NGP, pos, number, BoxSize, None, 2)


def NGPWc2D(pos, number, W, BoxSize, threads):
    _ckernel(N.NGP, pos, number, BoxSize, W, 2)


def CICc3D(pos, number, BoxSize, threads):
    _ckernel(N.CIC, pos, number, BoxSize, None, 3)


def CICWc3D(pos, number, W, BoxSize, threads):
    _ckernel(N.CIC, pos, number, BoxSize, W, 3)


def CICc2D(pos, number, BoxSize, threads):
    _ckernel(N.CIC, pos, number, BoxSize, None, 2)


def CICWc2D(pos, number, W, BoxSize, threads):
    _ckernel(N.CIC, pos, number, BoxSize, W, 2)


def TSCc3D(pos, number, BoxSize, threads):
    _ckernel(N.TSC, pos, number, BoxSize, None, 3)


def TSCWc3D(pos, number, W, BoxSize, threads):
    _ckernel(N.TSC, pos, number, BoxSize, W, 3)


def TSCc2D(pos, number, BoxSize, threads):
    _ckernel(N.TSC, pos, number, BoxSize, None, 2)


def TSCWc2D(pos, number, W, BoxSize, threads):
    _ckernel(N.TSC, pos, number, BoxSize, W, 2)


def PCSc3D(pos, number, BoxSize, threads):
    _ckernel(N.PCS, pos, number, BoxSize, None, 3)


def PCSWc3D(pos, number, W, BoxSize, threads):
    _ckernel(N.PCS, pos, number, BoxSize, W, 3)


def PCSc2D(pos, number, BoxSize, threads):
    _ckernel(N.PCS, pos, number, BoxSize, None, 2)


def PCSWc2D(pos, number, W, BoxSize, threads):
    _ckernel(N.PCS, pos, number, BoxSize, W, 2)


def CIC_interp(density, BoxSize, pos, den):
    """CIC-interpolate a 3D float32 grid at positions, writing den in place
    (MAS_library.pyx:558-599)."""
    _check_f32(density, 3, 'density')
    _check_f32(pos, 2, 'pos')
    _check_f32(den, 1, 'den')
    if N.is_device_array(density):
        ctx = N.default_context()
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_cic_interp(ctx.handle, N.device_ptr(density), density.shape[0],
                                           float(BoxSize), N.device_ptr(pos), pos.shape[0],
                                           N.device_ptr(den), N.PTR_DEVICE))
        return
    ctx = N.default_context()
    d = np.ascontiguousarray(density)
    p = np.ascontiguousarray(pos)
    out = den if den.flags['C_CONTIGUOUS'] else np.empty_like(den)
    N.check(N.lib().pyl_cic_interp(ctx.handle, d.ctypes.data, d.shape[0], float(BoxSize),
                                   p.ctypes.data, p.shape[0], out.ctypes.data, 0))
    if out is not den:
        den[...] = out


def MA_atomic(pos, number, BoxSize, MAS='CIC', W=None):
    """The plain GPU form: per-particle global float atomics (not bitwise reproducible).
    Same semantics as MA for 3D grids; used as the comparison baseline."""
    _check_f32(pos, 2, 'pos')
    _check_f32(number, pos.shape[1], 'number')
    flags = N.MA_ATOMIC | (N.MA_CYTHON_2D | N.MA_RENORM_2D if pos.shape[1] == 2 else 0)
    _deposit(pos, number, BoxSize, _MAS_IDS[MAS], W, flags)


def MA_rsd(pos, vel, number, BoxSize, Hubble, redshift, axis, MAS='CIC', W=None, overwrite=False,
           scale=1.0, shift=0.0):
    """MA of the redshift-space positions pos[:,axis] + vel[:,axis]*(1+z)/H (wrapped as
    redshift_space_library.pyx:40-46) without modifying pos: the pos_redshift_space
    + MA sequence of the reference pipelines in one pass over the particles.

    3D only.  number += deposit, or with overwrite=True number = deposit/scale - shift
    (scale = mean density and shift = 1 give the density contrast directly)."""
    _check_f32(pos, 2, 'pos')
    _check_f32(vel, 2, 'vel')
    _check_f32(number, 3, 'number')
    if pos.shape[1] != 3 or vel.shape != pos.shape:
        raise ValueError('pos and vel must both be (N, 3)')
    if W is not None:
        _check_f32(W, 1, 'W')
    if MAS not in _MAS_IDS:
        print('option not valid!!!')
        sys.exit()
    flags = N.MA_OVERWRITE if overwrite else 0
    dims, n = number.shape[0], pos.shape[0]
    ctx = N.default_context()
    arrs = [pos, vel, number] + ([] if W is None else [W])
    dev = [N.is_device_array(a) for a in arrs]
    if any(dev) and not all(dev):
        raise ValueError('pos, vel, number and W must all be on the GPU or all on the host')
    if all(dev):
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_ma_rsd(ctx.handle, N.device_ptr(pos), N.device_ptr(vel), n,
                                       None if W is None else N.device_ptr(W), N.device_ptr(number),
                                       dims, float(BoxSize), _MAS_IDS[MAS], float(Hubble),
                                       float(redshift), int(axis), float(scale), float(shift),
                                       flags | N.PTR_DEVICE))
        return
    p, v = np.ascontiguousarray(pos), np.ascontiguousarray(vel)
    w = None if W is None else np.ascontiguousarray(W)
    if not number.flags['C_CONTIGUOUS']:
        raise ValueError('number must be C-contiguous')
    N.check(N.lib().pyl_ma_rsd(ctx.handle, p.ctypes.data, v.ctypes.data, n,
                               None if w is None else w.ctypes.data, number.ctypes.data, dims,
                               float(BoxSize), _MAS_IDS[MAS], float(Hubble), float(redshift),
                               int(axis), float(scale), float(shift), flags))
