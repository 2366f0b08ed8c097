"""Drop-in redshift_space_library (library/redshift_space_library/redshift_space_library.pyx).

    pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis)
        pos[:, axis] += vel[:, axis] * (1 + z)/H, then the periodic wrap of
        redshift_space_library.pyx:40-46; pos is modified in place.

pos and vel are (N, 3) float32: C-contiguous numpy arrays (staged through the
device) or contiguous torch tensors on the GPU.  The same shift can be fused into
the mass assignment without touching pos (MAS_library.MA_rsd / pyl_ma_rsd).
"""
import numpy as np

from . import _native as N


def _check(a, name):
    if N.is_device_array(a):
        import torch
        if a.dtype != torch.float32 or a.dim() != 2 or a.shape[1] != 3 or not a.is_contiguous():
            raise ValueError('%s must be a contiguous (N, 3) float32 tensor' % name)
        return True
    if not isinstance(a, np.ndarray) or a.dtype != np.float32:
        raise ValueError("Buffer dtype mismatch, expected 'float' for %s" % name)
    if a.ndim != 2 or a.shape[1] != 3 or not a.flags['C_CONTIGUOUS']:
        raise ValueError('%s must be a C-contiguous (N, 3) array' % name)
    return False


def pos_redshift_space(pos, vel, BoxSize, Hubble, redshift, axis):
    dev = _check(pos, 'pos')
    if _check(vel, 'vel') != dev:
        raise ValueError('pos and vel must both be on the GPU or both on the host')
    if vel.shape[0] != pos.shape[0]:
        raise ValueError('pos and vel have different lengths')
    ctx = N.default_context()
    if dev:
        with N.torch_stream(ctx, True):
            N.check(N.lib().pyl_pos_redshift_space(ctx.handle, N.device_ptr(pos), N.device_ptr(vel),
                                                   pos.shape[0], float(BoxSize), float(Hubble),
                                                   float(redshift), int(axis), N.PTR_DEVICE))
        return
    N.check(N.lib().pyl_pos_redshift_space(ctx.handle, pos.ctypes.data, vel.ctypes.data,
                                           pos.shape[0], float(BoxSize), float(Hubble),
                                           float(redshift), int(axis), 0))
