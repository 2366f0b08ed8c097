"""Writes the Gadget reader fixtures: tests/golden/gadget/*  and  tests/golden/gadget_golden.npz.

Run in a container that holds the reference sources (read-only):

    python tests/golden/make_gadget_golden.py [/root/reference]

Small binary snapshots are written by pylians3_amd.readgadget.write_snapshot
(format I and II, both byte orders, one and several files, a header mass for
one type), and then read by the reference's own reader, library/readsnap.py
(snapshot_header :26-95, read_block :161-381), imported from the reference
tree.  Its header fields and blocks are stored as the expected values: the
committed snapshots plus these vectors pin pylians3_amd.readgadget against the
reference reader without the reference at test time.
"""
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = [  # name, format, endian, files, redshift, massarr
    ('f1_le_1', 1, '<', 1, 0.0, [0.0, 0.25, 0.0, 0.0, 0.0, 0.0]),
    ('f2_le_3', 2, '<', 3, 1.0, [0.0, 0.5, 0.0, 0.0, 0.0, 0.0]),
    ('f1_be_2', 1, '>', 2, 3.0, [0.0, 0.0, 0.0, 0.0, 0.0, 0.0]),
    ('f2_be_1', 2, '>', 1, 0.5, [0.0, 0.125, 0.0, 0.0, 0.0, 0.0]),
]
NTYPE = (0, 180, 90, 0, 25, 0)
BLOCKS = ('POS ', 'VEL ', 'ID  ', 'MASS')
HEADER = ('npart', 'massarr', 'time', 'redshift', 'nall', 'filenum', 'boxsize', 'omega_m',
          'omega_l', 'hubble', 'format')


def write_cases(outdir):
    from pylians3_amd import readgadget as RG
    os.makedirs(outdir, exist_ok=True)
    for ci, (name, fmt, endian, nfiles, z, massarr) in enumerate(CASES):
        rng = np.random.default_rng(100 + ci)
        pos = [(rng.random((k, 3)) * 50000.0).astype(np.float32) for k in NTYPE]
        vel = [rng.normal(0, 300, (k, 3)).astype(np.float32) for k in NTYPE]
        mass = [(rng.random(k) + 0.1).astype(np.float32) for k in NTYPE]
        RG.write_snapshot(os.path.join(outdir, name), pos, 50000.0, redshift=z, vel_by_type=vel,
                          mass_by_type=mass, massarr=massarr, nfiles=nfiles, fmt=fmt,
                          endian=endian)


def main(ref_root):
    outdir = os.path.join(HERE, 'gadget')
    write_cases(outdir)
    spec = importlib.util.spec_from_file_location(
        'readsnap', os.path.join(ref_root, 'library', 'readsnap.py'))
    rs = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(rs)
    out = {}
    for name, fmt, endian, nfiles, z, massarr in CASES:
        snap = os.path.join(outdir, name)
        first = snap if nfiles == 1 else snap + '.0'
        h = rs.snapshot_header(first)
        for k in HEADER:
            out['%s/%s' % (name, k)] = np.asarray(getattr(h, k))
        for t in (1, 2, 4):
            for b in BLOCKS:
                out['%s/%s/%d' % (name, b.strip(), t)] = np.asarray(rs.read_block(snap, b, parttype=t))
    np.savez_compressed(os.path.join(HERE, 'gadget_golden.npz'), **out)
    print('wrote', len(out), 'arrays for', len(CASES), 'snapshots')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
