"""Gadget snapshot reader (binary format I and II), drop-in for the parts of
library/readgadget.py and library/readsnap.py that the density-field pipeline uses.

    header(snapshot)                 readgadget.py:21-83  (binary branch)
        .time .redshift .npart .nall .filenum .massarr .boxsize .omega_m
        .omega_l .hubble .cooling .format .Hubble
    read_field(snapshot, block, ptype)   readgadget.py:86-122 (one file)
    read_block(snapshot, block, ptype)   readgadget.py:126-178 (all files, list of types)
    write_snapshot(...)              writes format I/II files (tests, synthetic inputs)

Blocks: "POS ", "VEL ", "ID  ", "MASS".  Layout of a binary snapshot (the Gadget-2
convention readsnap.py follows): Fortran-style records (int32 byte count, payload,
int32 byte count); a 256-byte header (npart[6] int32, massarr[6] float64, time,
redshift float64, sfr, feedback int32, nall[6] uint32, cooling, filenum int32,
boxsize, omega_m, omega_l, hubble float64, padding); then POS and VEL (float32 x3
per particle, ordered by type), ID (uint32, or uint64 when the record is twice as
long), MASS (float32, only for types whose massarr entry is 0).  Format II adds an
8-byte record with the 4-character block name before every block.  Files of either
endianness are read; VEL is multiplied by sqrt(time) when the redshift is not zero
(readsnap.py physical_velocities).  HDF5 snapshots need h5py, which this package
does not depend on: they raise NotImplementedError.

Blocks are memory-mapped, so reading a field of a large snapshot touches only
that field's bytes.
"""
import math
import os

import numpy as np

_HEADER_FIELDS = [('npart', '<i4', 6), ('massarr', '<f8', 6), ('time', '<f8', 1),
                  ('redshift', '<f8', 1), ('sfr', '<i4', 1), ('feedback', '<i4', 1),
                  ('nall', '<u4', 6), ('cooling', '<i4', 1), ('filenum', '<i4', 1),
                  ('boxsize', '<f8', 1), ('omega_m', '<f8', 1), ('omega_l', '<f8', 1),
                  ('hubble', '<f8', 1)]
_BLOCK_ORDER = ['HEAD', 'POS ', 'VEL ', 'ID  ', 'MASS']   # format I block numbers 1..5


def fname_format(snapshot):
    """readgadget.py:7-18: resolve the first file of a snapshot and its format."""
    if os.path.exists(snapshot):
        if snapshot[-4:] == 'hdf5':
            return snapshot, 'hdf5'
        return snapshot, 'binary'
    if os.path.exists(snapshot + '.0'):
        return snapshot + '.0', 'binary'
    if os.path.exists(snapshot + '.hdf5'):
        return snapshot + '.hdf5', 'hdf5'
    if os.path.exists(snapshot + '.0.hdf5'):
        return snapshot + '.0.hdf5', 'hdf5'
    raise Exception('File not found!')


def _no_hdf5():
    raise NotImplementedError('HDF5 snapshots need h5py, which is not available here')


def _endian(f):
    first = np.frombuffer(f.read(4), '<i4')[0]
    if first in (8, 256):
        return '<', (2 if first == 8 else 1)
    first = np.frombuffer(first.tobytes(), '>i4')[0]
    if first in (8, 256):
        return '>', (2 if first == 8 else 1)
    raise ValueError('incorrect file format encountered when reading header')


class _FileLayout:
    """Header and block table (name -> (payload offset, payload bytes)) of one file."""

    def __init__(self, filename):
        self.filename = filename
        with open(filename, 'rb') as f:
            self.order, self.format = _endian(f)
            f.seek(0)
            size = os.path.getsize(filename)
            blocks = {}
            num = 0
            pos = 0
            while pos < size:
                name = None
                if self.format == 2:
                    f.seek(pos + 4)
                    name = f.read(4).decode('ascii', 'replace')
                    pos += 16
                f.seek(pos)
                nbytes = int(np.frombuffer(f.read(4), self.order + 'u4')[0])
                num += 1
                if name is None:
                    name = _BLOCK_ORDER[num - 1] if num <= len(_BLOCK_ORDER) else 'B%03d' % num
                blocks[name] = (pos + 4, nbytes)
                f.seek(pos + 4 + nbytes)
                check = int(np.frombuffer(f.read(4), self.order + 'u4')[0])
                if check != nbytes:
                    raise ValueError('something wrong: block %r of %s has inconsistent record '
                                     'markers' % (name, filename))
                pos += 8 + nbytes
            self.blocks = blocks
            off, nbytes = blocks['HEAD']
            f.seek(off)
            raw = f.read(256)
        h = {}
        p = 0
        for name, dt, count in _HEADER_FIELDS:
            dt = np.dtype(dt).newbyteorder(self.order)
            v = np.frombuffer(raw, dt, count=count, offset=p)
            p += dt.itemsize * count
            h[name] = v.astype(v.dtype.newbyteorder('=')) if count > 1 else v[0].item()
        self.h = h


class header:
    """readgadget.py:21-83 for binary snapshots."""

    def __init__(self, snapshot):
        filename, fformat = fname_format(snapshot)
        if fformat == 'hdf5':
            _no_hdf5()
        L = _FileLayout(filename)
        h = L.h
        self.time = h['time']
        self.redshift = h['redshift']
        self.boxsize = h['boxsize']
        self.filenum = h['filenum']
        self.omega_m = h['omega_m']
        self.omega_l = h['omega_l']
        self.hubble = h['hubble']
        self.massarr = h['massarr'].astype(np.float64)
        self.npart = h['npart'].astype(np.int32)
        self.nall = h['nall'].astype(np.uint32)
        self.cooling = h['cooling']
        self.format = L.format
        # km/s/(Mpc/h)
        self.Hubble = 100.0 * np.sqrt(self.omega_m * (1.0 + self.redshift) ** 3 + self.omega_l)


def _block_dtype(block):
    if block in ('POS ', 'VEL '):
        return np.dtype((np.float32, 3))
    if block == 'MASS':
        return np.dtype(np.float32)
    if block == 'ID  ':
        return np.dtype(np.uint32)
    raise Exception('block not implemented in readgadget!')


def _read_file_block(filename, block, ptype):
    """One particle type's slice of a block of one file (readsnap.py:186-371, single file)."""
    L = _FileLayout(filename)
    h = L.h
    npart = h['npart'].astype(np.int64)
    massarr = h['massarr']
    if block == 'MASS' and massarr[ptype] > 0:
        return np.ones(npart[ptype], np.float32) * np.float32(massarr[ptype])
    dt = _block_dtype(block)
    if block == 'MASS':
        present = massarr == 0
    else:
        present = np.ones(6, bool)
    if not present[ptype]:
        raise Exception('no data for particle type %d in block %s' % (ptype, block))
    if block not in L.blocks:
        raise Exception('block %r not found in %s' % (block, filename))
    off, nbytes = L.blocks[block]
    nrec = int(np.sum(npart[present]))
    if block == 'ID  ' and nbytes == 8 * nrec:
        dt = np.dtype(np.uint64)
    if dt.itemsize * nrec != nbytes:
        raise Exception('something wrong with blocksize! expected = %d actual = %d'
                        % (dt.itemsize * nrec, nbytes))
    first = int(np.sum(npart[:ptype][present[:ptype]]))
    base = dt.base.newbyteorder(L.order)
    shape = (int(npart[ptype]),) + dt.shape
    mm = np.memmap(filename, dtype=base, mode='r', offset=off + first * dt.itemsize, shape=shape)
    out = np.array(mm, dtype=base.newbyteorder('='))
    del mm
    if block == 'VEL ' and h['redshift'] != 0:
        out *= np.float32(math.sqrt(h['time']))
    return out


def read_field(snapshot, block, ptype):
    """readgadget.py:86-122: block of one particle type from one file."""
    filename, fformat = fname_format(snapshot)
    if fformat == 'hdf5':
        _no_hdf5()
    return _read_file_block(filename, block, ptype)


def subfiles(snapshot):
    """File names of a (possibly multi-file) binary snapshot, in order."""
    filename, fformat = fname_format(snapshot)
    if fformat == 'hdf5':
        _no_hdf5()
    n = header(filename).filenum
    if os.path.exists(snapshot) and n == 1:
        return [snapshot]
    return [snapshot + '.%d' % i for i in range(n)] if n > 1 else [filename]


def read_block(snapshot, block, ptype, verbose=False):
    """readgadget.py:126-178: a block over all files, for a list of particle types
    (concatenated by type, then by file)."""
    parts = []
    for pt in ptype:
        for fn in subfiles(snapshot):
            if _FileLayout(fn).h['npart'][pt] > 0:
                parts.append(_read_file_block(fn, block, pt))
    head = header(fname_format(snapshot)[0])
    ntot = int(sum(int(head.nall[i]) for i in ptype))
    if not parts:
        return np.zeros((0,) + _block_dtype(block).shape, _block_dtype(block).base)
    out = np.concatenate(parts)
    if out.shape[0] != ntot:
        raise Exception('not all particles read!!!!')
    return out


def write_snapshot(snapshot, pos_by_type, BoxSize, redshift=0.0, omega_m=0.3, omega_l=0.7,
                   hubble=0.7, vel_by_type=None, mass_by_type=None, massarr=None, nfiles=1,
                   fmt=1, endian='<', ids64=False):
    """Write a binary Gadget snapshot (format 1 or 2) of up to 6 particle types.

    pos_by_type[t]: (n_t, 3) float32 in kpc/h (the reader's unit convention); vel
    likewise (km/s, stored as v/sqrt(a)); mass_by_type[t] (n_t,) in 1e10 Msun/h for
    the types whose massarr entry is 0.  Particles of each type are split evenly
    over nfiles files (snapshot.0, snapshot.1, ... when nfiles > 1)."""
    ntypes = 6
    pos_by_type = list(pos_by_type) + [np.zeros((0, 3), np.float32)] * (ntypes - len(pos_by_type))
    if vel_by_type is None:
        vel_by_type = [np.zeros_like(p) for p in pos_by_type]
    else:
        vel_by_type = list(vel_by_type) + [np.zeros((0, 3), np.float32)] * (ntypes - len(vel_by_type))
    massarr = np.zeros(6) if massarr is None else np.asarray(massarr, np.float64)
    if mass_by_type is None:
        mass_by_type = [np.zeros(len(p), np.float32) for p in pos_by_type]
    nall = np.array([len(p) for p in pos_by_type], np.int64)
    a = 1.0 / (1.0 + redshift)
    vfac = math.sqrt(a) if redshift != 0 else 1.0
    splits = [np.array_split(np.arange(n), nfiles) for n in nall]
    u4 = np.dtype(endian + 'u4')
    for fi in range(nfiles):
        fn = snapshot if nfiles == 1 else '%s.%d' % (snapshot, fi)
        idx = [splits[t][fi] for t in range(ntypes)]
        npart = np.array([len(i) for i in idx], np.int32)
        head = bytearray(256)
        vals = dict(npart=npart, massarr=massarr, time=a, redshift=redshift, sfr=0, feedback=0,
                    nall=nall.astype(np.uint32), cooling=0, filenum=nfiles, boxsize=BoxSize,
                    omega_m=omega_m, omega_l=omega_l, hubble=hubble)
        p = 0
        for name, dt, count in _HEADER_FIELDS:
            d = np.dtype(dt).newbyteorder(endian)
            b = np.asarray(vals[name], dtype=d).reshape(count).tobytes()
            head[p:p + len(b)] = b
            p += len(b)
        id0 = 0
        ids = []
        for t in range(ntypes):
            start = int(np.sum(nall[:t])) + (idx[t][0] if len(idx[t]) else 0)
            ids.append(np.arange(start, start + len(idx[t])))
            id0 += len(idx[t])
        blocks = [('HEAD', bytes(head)),
                  ('POS ', np.concatenate([pos_by_type[t][idx[t]] for t in range(ntypes)])
                   .astype(endian + 'f4').tobytes()),
                  ('VEL ', np.concatenate([vel_by_type[t][idx[t]] / np.float32(vfac)
                                           for t in range(ntypes)]).astype(endian + 'f4').tobytes()),
                  ('ID  ', np.concatenate(ids).astype(endian + ('u8' if ids64 else 'u4')).tobytes())]
        mblk = [mass_by_type[t][idx[t]] for t in range(ntypes) if massarr[t] == 0 and npart[t] > 0]
        if mblk:
            blocks.append(('MASS', np.concatenate(mblk).astype(endian + 'f4').tobytes()))
        with open(fn, 'wb') as f:
            for name, payload in blocks:
                if fmt == 2:
                    f.write(np.array([8], u4).tobytes())
                    f.write(name.encode('ascii'))
                    f.write(np.array([len(payload) + 8, 8], u4).tobytes())
                f.write(np.array([len(payload)], u4).tobytes())
                f.write(payload)
                f.write(np.array([len(payload)], u4).tobytes())
