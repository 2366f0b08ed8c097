"""Batched MA and Pk/XPk over many realisations (one launch sequence per batch).

    MA_batch(pos, offsets, grids, BoxSize, MAS='CIC', W=None, accumulate=True)
        particles of realisation r are pos[offsets[r]:offsets[r+1]]; grids is
        (nreal, d, d, d) float32; same result per grid as MAS_library.MA on it.
    Pk_batch(deltas, BoxSize, axis=2, MAS=('CIC',), scales=None, shifts=None)
        deltas is a list of nfields (nreal, d, d, d) float32 arrays (or one
        (nreal, nfields, d, d, d) array); returns one result per
        realisation with the attributes of Pk_library.Pk (nfields == 1) or
        Pk_library.XPk (nfields >= 2).  The optional prologue applies
        delta = delta/scale - shift per field in float32 before the FFT.

Arrays may be numpy (staged) or contiguous torch tensors on the GPU.  Each
realisation's result is bitwise identical to the corresponding single call:
the tiled scatter and the binning reduce in an order that does not depend on
the batch.
"""
import ctypes

import numpy as np

from . import _native as N
from .MAS_library import _MAS_IDS
from .Pk_library import MAS_function


class Result:
    """Attributes of Pk_library.Pk / XPk for one realisation."""


def MA_batch(pos, offsets, grids, BoxSize, MAS='CIC', W=None, accumulate=True, atomic=False,
             ctx=None):
    ctx = ctx or N.default_context()
    offs = np.ascontiguousarray(offsets, dtype=np.int64)
    nreal = len(offs) - 1
    dims = grids.shape[1]
    axes = pos.shape[1]
    if grids.shape[0] != nreal:
        raise ValueError('grids must have one grid per realisation')
    flags = 0 if accumulate else N.MA_OVERWRITE
    if atomic:
        flags |= N.MA_ATOMIC
    if N.is_device_array(pos):
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_ma_batch(ctx.handle, N.device_ptr(pos), offs.ctypes.data, nreal, axes,
                                         None if W is None else N.device_ptr(W),
                                         N.device_ptr(grids), dims, float(BoxSize), _MAS_IDS[MAS],
                                         flags | N.PTR_DEVICE))
        return grids
    p = np.ascontiguousarray(pos, dtype=np.float32)
    w = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
    if not grids.flags['C_CONTIGUOUS'] or grids.dtype != np.float32:
        raise ValueError('grids must be a C-contiguous float32 array')
    N.check(N.lib().pyl_ma_batch(ctx.handle, p.ctypes.data, offs.ctypes.data, nreal, axes,
                                 None if w is None else w.ctypes.data, grids.ctypes.data, dims,
                                 float(BoxSize), _MAS_IDS[MAS], flags))
    return grids


def _fields(deltas):
    """List of F field arrays, each (nreal, d, d, d) contiguous float32.  Accepts such
    a list, or one (nreal, F, d, d, d) array (split into per-field copies)."""
    if isinstance(deltas, (list, tuple)):
        fields = list(deltas)
    else:
        fields = [deltas[:, f] for f in range(deltas.shape[1])]
    out = []
    for a in fields:
        if N.is_device_array(a):
            out.append(a.contiguous())
        else:
            out.append(np.ascontiguousarray(a, dtype=np.float32))
    return out


def Pk_batch(deltas, BoxSize, axis=2, MAS=('CIC',), scales=None, shifts=None, destroy=False,
             ctx=None):
    ctx = ctx or N.default_context()
    fields = _fields(deltas)
    F = len(fields)
    nreal, dims = fields[0].shape[0], fields[0].shape[1]
    if len(MAS) != F:
        raise ValueError('one MAS per field')
    for a in fields:
        if tuple(a.shape) != (nreal, dims, dims, dims):
            raise ValueError('every field must be (nreal, dims, dims, dims)')
    mas = (ctypes.c_int * F)(*[MAS_function(m) for m in MAS])
    allocs = [N.alloc_pk_out(dims, F) for _ in range(nreal)]
    outs = (N.PkOut * nreal)(*[a[0] for a in allocs])
    sc = None if scales is None else np.ascontiguousarray(scales, dtype=np.float32).reshape(-1)
    sh = None if shifts is None else np.ascontiguousarray(shifts, dtype=np.float32).reshape(-1)
    for a in (sc, sh):
        if a is not None and a.size != nreal * F:
            raise ValueError('scales/shifts need nreal*nfields values')
    dev = N.is_device_array(fields[0])
    ptrs = (ctypes.c_void_p * F)(*[(N.device_ptr(a) if dev else a.ctypes.data) for a in fields])
    flags = (N.PTR_DEVICE | (N.PK_DESTROY_INPUT if destroy else 0)) if dev else 0
    with N.torch_stream(ctx, dev):
        N.check(N.lib().pyl_pk_batch(ctx.handle, ptrs, nreal, F, mas, dims, float(BoxSize),
                                     int(axis), None if sc is None else sc.ctypes.data,
                                     None if sh is None else sh.ctypes.data, flags, outs))
    X = F * (F - 1) // 2
    results = []
    for _, arrs, S in allocs:
        r = Result()
        r.k1D, r.Nmodes1D = arrs['k1D'], arrs['Nmodes1D']
        r.kpar, r.kper, r.Nmodes2D = arrs['kpar'], arrs['kper'], arrs['Nmodes2D']
        r.k3D, r.Nmodes3D = arrs['k3D'], arrs['Nmodes3D']
        if F == 1:
            r.Pk = arrs['Pk3D'][:, :, 0]
            r.Pk1D = arrs['Pk1D'][:, 0]
            r.Pk2D = arrs['Pk2D'][:, 0]
            r.Pkphase = arrs['Pkphase']
        else:
            r.Pk, r.XPk = arrs['Pk3D'][:, :, :F], arrs['Pk3D'][:, :, F:F + X]
            r.Pk1D, r.PkX1D = arrs['Pk1D'][:, :F], arrs['Pk1D'][:, F:F + X]
            r.Pk2D, r.PkX2D = arrs['Pk2D'][:, :F], arrs['Pk2D'][:, F:F + X]
        results.append(r)
    return results
