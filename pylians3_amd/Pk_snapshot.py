"""Power spectra of Gadget snapshots on the GPU (library/Pk_library/Pk_snapshot.py).

    Pk_comp(snapshot_fname, ptype, dims, do_RSD, axis, cpus, folder_out)    :34-91
    Pk_Gadget(snapshot_fname, dims, particle_type, do_RSD, axis, cpus,
              folder_out=None)                                               :107-268

Same inputs, printed summary and output files (k, P0, P2, P4, Nmodes columns,
names from name_dict) as the reference.  The particles go to the GPU once; the
redshift-space shift, the CIC assignment and delta = rho/mean - 1 are one fused
pass (pyl_ma_rsd / pyl_ma_delta), and the spectra run on the device field.

Differences kept on purpose: ptype = -1 reads every particle type (the
reference's read_block([-1]) counts only type 5); delta /= mean is a float32
division by the float32 mean in every case.
"""
import os

import numpy as np

from . import _native as N
from . import Pk_library as PKL
from . import readgadget
from .MAS_library import _MAS_IDS

rho_crit = 2.77536627e11   # h^2 Msun/Mpc^3 (units_library.py:10)

name_dict = {'0': 'GAS', '01': 'GCDM', '02': 'GNU', '04': 'Gstars',
             '1': 'CDM', '12': 'CDMNU', '14': 'CDMStars',
             '2': 'NU', '24': 'NUStars',
             '4': 'Stars',
             '-1': 'matter'}


def _delta_device(pos, vel, mass, dims, BoxSize, mean, do_RSD, Hubble, redshift, axis):
    """delta = CIC(pos [+RSD], W=mass)/mean - 1 on the device, one fused pass."""
    import torch
    ctx = N.default_context()
    delta = torch.empty((dims,) * 3, dtype=torch.float32, device='cuda')
    tp = torch.from_numpy(np.ascontiguousarray(pos, np.float32)).cuda()
    tw = None if mass is None else torch.from_numpy(np.ascontiguousarray(mass, np.float32)).cuda()
    scale = float(np.float32(mean))
    with N.torch_stream(ctx, True):
        if do_RSD:
            tv = torch.from_numpy(np.ascontiguousarray(vel, np.float32)).cuda()
            N.check(N.lib().pyl_ma_rsd(ctx.handle, N.device_ptr(tp), N.device_ptr(tv), tp.shape[0],
                                       None if tw is None else N.device_ptr(tw), N.device_ptr(delta),
                                       dims, float(BoxSize), _MAS_IDS['CIC'], float(Hubble),
                                       float(redshift), int(axis), scale, 1.0,
                                       N.PTR_DEVICE | N.MA_OVERWRITE))
        else:
            N.check(N.lib().pyl_ma_delta(ctx.handle, N.device_ptr(tp), tp.shape[0], 3,
                                         None if tw is None else N.device_ptr(tw), N.device_ptr(delta),
                                         dims, float(BoxSize), _MAS_IDS['CIC'], scale, 1.0,
                                         N.PTR_DEVICE))
    return delta


def _types(ptype):
    return [0, 1, 2, 3, 4, 5] if ptype == -1 else [ptype]


def Pk_comp(snapshot_fname, ptype, dims, do_RSD, axis, cpus, folder_out):
    print('Computing power spectrum...')
    head = readgadget.header(snapshot_fname)
    BoxSize = head.boxsize / 1e3
    Masses = head.massarr * 1e10
    Nall = head.nall.astype(np.int64)
    Ntotal = np.sum(Nall, dtype=np.int64)
    Omega_m = head.omega_m
    Omega_l = head.omega_l
    redshift = head.redshift
    Hubble = 100.0 * np.sqrt(Omega_m * (1.0 + redshift) ** 3 + Omega_l)
    z = '%.3f' % redshift
    fout = folder_out + '/Pk_' + name_dict[str(ptype)]
    if do_RSD:
        fout += ('_RS_axis=' + str(axis) + '_z=' + z + '.dat')
    else:
        fout += ('_z=' + z + '.dat')
    pos = readgadget.read_block(snapshot_fname, 'POS ', _types(ptype)) / 1e3
    print('%.3f < X [Mpc/h] < %.3f' % (np.min(pos[:, 0]), np.max(pos[:, 0])))
    print('%.3f < Y [Mpc/h] < %.3f' % (np.min(pos[:, 1]), np.max(pos[:, 1])))
    print('%.3f < Z [Mpc/h] < %.3f\n' % (np.min(pos[:, 2]), np.max(pos[:, 2])))
    vel = None
    if do_RSD:
        print('moving particles to redshift-space...')
        vel = readgadget.read_block(snapshot_fname, 'VEL ', _types(ptype))
        print('done')
    M = None
    if ptype == -1:
        if Nall[0] == 0:
            M = np.zeros(Ntotal, dtype=np.float32)
            offset = 0
            for pt in [0, 1, 2, 3, 4, 5]:
                M[offset:offset + Nall[pt]] = Masses[pt]
                offset += Nall[pt]
        else:
            M = (readgadget.read_block(snapshot_fname, 'MASS', [0, 1, 2, 3, 4, 5]) * 1e10)
        mean = np.sum(M, dtype=np.float64) / dims ** 3
    else:
        mean = len(pos) * 1.0 / dims ** 3
    delta = _delta_device(pos, vel, M, dims, BoxSize, mean, do_RSD, Hubble, redshift, axis)
    del pos, vel, M
    Pk = PKL.Pk(delta, BoxSize, axis=axis, MAS='CIC', threads=cpus)
    del delta
    np.savetxt(fout, np.transpose([Pk.k3D, Pk.Pk[:, 0], Pk.Pk[:, 1], Pk.Pk[:, 2], Pk.Nmodes3D]))
    return Pk


def Pk_Gadget(snapshot_fname, dims, particle_type, do_RSD, axis, cpus, folder_out=None):
    if folder_out is None:
        folder_out = os.getcwd()
    if len(particle_type) == 1:
        Pk_comp(snapshot_fname, particle_type[0], dims, do_RSD, axis, cpus, folder_out)
        return None
    print('\nREADING SNAPSHOTS PROPERTIES')
    head = readgadget.header(snapshot_fname)
    BoxSize = head.boxsize / 1e3
    Nall = head.nall.astype(np.int64)
    Masses = head.massarr * 1e10
    Omega_m = head.omega_m
    Omega_l = head.omega_l
    redshift = head.redshift
    Hubble = 100.0 * np.sqrt(Omega_m * (1.0 + redshift) ** 3 + Omega_l)
    z = '%.3f' % redshift
    dims3 = dims ** 3
    Omega_c = Masses[1] * Nall[1] / BoxSize ** 3 / rho_crit
    Omega_n = Masses[2] * Nall[2] / BoxSize ** 3 / rho_crit
    Omega_g, Omega_s = 0.0, 0.0
    if Nall[0] > 0:
        if Masses[0] > 0:
            Omega_g = Masses[0] * Nall[0] / BoxSize ** 3 / rho_crit
            Omega_s = Masses[4] * Nall[4] / BoxSize ** 3 / rho_crit
        else:
            mass = readgadget.read_block(snapshot_fname, 'MASS', [0]) * 1e10
            Omega_g = np.sum(mass, dtype=np.float64) / BoxSize ** 3 / rho_crit
            mass = readgadget.read_block(snapshot_fname, 'MASS', [4]) * 1e10
            Omega_s = np.sum(mass, dtype=np.float64) / BoxSize ** 3 / rho_crit
            del mass
    print('Omega_gas    = ', Omega_g)
    print('Omega_cdm    = ', Omega_c)
    print('Omega_nu     = ', Omega_n)
    print('Omega_star   = ', Omega_s)
    print('Omega_m      = ', Omega_g + Omega_c + Omega_n + Omega_s)
    print('Omega_m snap = ', Omega_m)
    Omega_dict = {0: Omega_g, 1: Omega_c, 2: Omega_n, 4: Omega_s}
    delta = [[], [], [], []]
    index_dict = {0: 0, 1: 1, 2: 2, 4: 3}
    if do_RSD:
        suffix = '_RS_axis=' + str(axis) + '_z=' + z + '.dat'
    else:
        suffix = '_z=' + z + '.dat'
    for ptype in particle_type:
        pos = readgadget.read_block(snapshot_fname, 'POS ', [ptype]) / 1e3
        vel = readgadget.read_block(snapshot_fname, 'VEL ', [ptype]) if do_RSD else None
        index = index_dict[ptype]
        mean_number = len(pos) * 1.0 / dims3
        delta[index] = _delta_device(pos, vel, None, dims, BoxSize, mean_number, do_RSD, Hubble,
                                     redshift, axis)
        del pos, vel
    for i, ptype1 in enumerate(particle_type):
        for ptype2 in particle_type[i + 1:]:
            index1 = index_dict[ptype1]
            index2 = index_dict[ptype2]
            fout1 = folder_out + '/Pk_' + name_dict[str(ptype1)] + suffix
            fout2 = folder_out + '/Pk_' + name_dict[str(ptype2)] + suffix
            fout12 = folder_out + '/Pk_' + name_dict[str(ptype1) + str(ptype2)] + suffix
            print('\nComputing the auto- and cross-power spectra of types: ', ptype1, '-', ptype2)
            print('saving results in:')
            print(fout1, '\n', fout2, '\n', fout12)
            data = PKL.XPk([delta[index1], delta[index2]], BoxSize, axis=axis, MAS=['CIC', 'CIC'],
                           threads=cpus)
            k = data.k3D
            Nmodes = data.Nmodes3D
            np.savetxt(fout12, np.transpose([k, data.XPk[:, 0, 0], data.XPk[:, 1, 0],
                                             data.XPk[:, 2, 0], Nmodes]))
            np.savetxt(fout1, np.transpose([k, data.Pk[:, 0, 0], data.Pk[:, 1, 0], data.Pk[:, 2, 0],
                                            Nmodes]))
            np.savetxt(fout2, np.transpose([k, data.Pk[:, 0, 1], data.Pk[:, 1, 1], data.Pk[:, 2, 1],
                                            Nmodes]))
    print('\ncomputing P(k) of all components')
    import torch
    delta_tot = torch.zeros((dims, dims, dims), dtype=torch.float32, device='cuda')
    Omega_tot = 0.0
    fout = folder_out + '/Pk_'
    for ptype in particle_type:
        index = index_dict[ptype]
        delta_tot += (float(Omega_dict[ptype]) * delta[index])
        Omega_tot += Omega_dict[ptype]
        fout += name_dict[str(ptype)] + '+'
    delta_tot /= float(Omega_tot)
    del delta
    fout = fout[:-1]
    data = PKL.Pk(delta_tot, BoxSize, axis=axis, MAS='CIC', threads=cpus)
    del delta_tot
    np.savetxt(fout + suffix, np.transpose([data.k3D, data.Pk[:, 0], data.Pk[:, 1], data.Pk[:, 2],
                                            data.Nmodes3D]))
    return None
