"""Density field of a Gadget snapshot on the GPU (library/MAS_library/MAS_gadget.py:7-83).

    density_field_gadget(snapshot_fname, ptypes, dims, MAS='CIC', do_RSD=False,
                         axis=0, verbose=True, device=False)

Same loop as the reference: over the sub-files of the snapshot and the requested
particle types, MA (+= onto one grid) of the positions in Mpc/h, moved to redshift
space along `axis` when do_RSD, weighted by the particle masses when several types
are combined.  Here the grid stays in GPU memory for the whole loop; each
sub-file's blocks are read (memory-mapped) on a host thread while the GPU deposits
the previous one, and the redshift-space shift is fused into the deposit
(pyl_ma_rsd).  Returns a float32 numpy array (or the device tensor with
device=True).

Deviation kept on purpose: for multi-type grids the reference builds the mass
array as float32 ones times a float64 header mass, which NumPy >= 2 promotes to
float64 (which MA then rejects); here the weights are float32, as under NumPy 1.
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _native as N
from . import readgadget
from .MAS_library import _MAS_IDS


def _load(snapshot, ptype, do_RSD, Masses, single_component):
    head = readgadget.header(snapshot)
    if head.npart[ptype] == 0:
        return None
    pos = readgadget.read_field(snapshot, 'POS ', ptype) / 1e3        # Mpc/h, float32
    vel = readgadget.read_field(snapshot, 'VEL ', ptype) if do_RSD else None
    if Masses[ptype] != 0:
        if single_component:
            mass = None
        else:
            mass = np.ones(head.npart[ptype], dtype=np.float32) * np.float32(Masses[ptype])
    else:
        mass = readgadget.read_field(snapshot, 'MASS', ptype) * 1e10
        mass = mass.astype(np.float32)
    return pos, vel, mass


def density_field_gadget(snapshot_fname, ptypes, dims, MAS='CIC', do_RSD=False, axis=0,
                         verbose=True, device=False):
    import torch
    start = time.time()
    if verbose:
        print('\nComputing density field of particles', ptypes)
    head = readgadget.header(snapshot_fname)
    BoxSize = head.boxsize / 1e3
    Masses = head.massarr * 1e10
    filenum = head.filenum
    redshift = head.redshift
    Hubble = head.Hubble
    if ptypes == [-1]:
        ptypes = [0, 1, 2, 3, 4, 5]
    single_component = len(ptypes) == 1
    density = torch.zeros((dims, dims, dims), dtype=torch.float32, device='cuda')
    ctx = N.default_context()
    mas = _MAS_IDS[MAS]
    work = []
    for i in range(filenum):
        snapshot = snapshot_fname if filenum == 1 else snapshot_fname + '.%d' % i
        for ptype in ptypes:
            work.append((snapshot, ptype))
    num = 0.0
    with ThreadPoolExecutor(1) as pool:
        nxt = pool.submit(_load, *work[0], do_RSD, Masses, single_component) if work else None
        for j in range(len(work)):
            item = nxt.result()
            nxt = (pool.submit(_load, *work[j + 1], do_RSD, Masses, single_component)
                   if j + 1 < len(work) else None)
            if item is None:
                continue
            pos, vel, mass = item
            tp = torch.from_numpy(pos).cuda()
            tv = torch.from_numpy(vel).cuda() if vel is not None else None
            tw = torch.from_numpy(mass).cuda() if mass is not None else None
            with N.torch_stream(ctx, True):
                if do_RSD:
                    N.check(N.lib().pyl_ma_rsd(ctx.handle, N.device_ptr(tp), N.device_ptr(tv),
                                               pos.shape[0], None if tw is None else N.device_ptr(tw),
                                               N.device_ptr(density), dims, float(BoxSize), mas,
                                               float(Hubble), float(redshift), int(axis), 1.0, 0.0,
                                               N.PTR_DEVICE))
                else:
                    N.check(N.lib().pyl_ma(ctx.handle, N.device_ptr(tp), pos.shape[0], 3,
                                           None if tw is None else N.device_ptr(tw),
                                           N.device_ptr(density), dims, float(BoxSize), mas,
                                           N.PTR_DEVICE))
            num += pos.shape[0] if mass is None else float(np.sum(mass, dtype=np.float64))
    if verbose:
        print('%.8e should be equal to\n%.8e' % (density.double().sum().item(), num))
        print('Time taken = %.2f seconds' % (time.time() - start))
    if device:
        return density
    return density.cpu().numpy()
