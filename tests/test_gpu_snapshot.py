"""Snapshot -> MA -> Pk pipeline on the GPU (MAS_gadget.density_field_gadget,
Pk_snapshot.Pk_comp / Pk_Gadget) against the oracle run on the same particles in
the reference's order: per sub-file and particle type, pos/1e3, the
redshift-space shift, then the serial float32 MA accumulating into one grid; the
density contrast and the oracle Pk for the spectra.  Grids per cell within 1e-5
(the MA tolerance), spectra per bin within 1e-5 of the monopole."""
import os

import numpy as np
import pytest

from conftest import assert_bins, assert_cells, contrib_counts

pytestmark = pytest.mark.gpu


def _snapshot(tmp_path, nfiles=3, fmt=2, redshift=0.5, seed=0):
    from pylians3_amd import readgadget as RG
    rng = np.random.default_rng(seed)
    box_kpc = 200000.0
    n = [0, 40000, 20000, 0, 5000, 0]
    pos = [(rng.random((k, 3)) * box_kpc).astype(np.float32) for k in n]
    vel = [rng.normal(0, 400, (k, 3)).astype(np.float32) for k in n]
    mass = [(rng.random(k) * 0.01).astype(np.float32) for k in n]
    snap = str(tmp_path / 'snapdir' / 'snap_002')
    os.makedirs(os.path.dirname(snap), exist_ok=True)
    RG.write_snapshot(snap, pos, box_kpc, redshift=redshift, vel_by_type=vel, mass_by_type=mass,
                      massarr=[0, 0.8, 0.1, 0, 0, 0], nfiles=nfiles, fmt=fmt)
    return snap


def _oracle_density(oracle, snap, ptypes, dims, do_RSD, axis, MAS='CIC'):
    from pylians3_amd import readgadget as RG
    h = RG.header(snap)
    box = h.boxsize / 1e3
    Masses = h.massarr * 1e10
    single = len(ptypes) == 1
    grid = np.zeros((dims,) * 3, np.float32)
    allpos = []
    for i in range(h.filenum):
        fn = snap if h.filenum == 1 else snap + '.%d' % i
        hh = RG.header(fn)
        for pt in ptypes:
            if hh.npart[pt] == 0:
                continue
            pos = RG.read_field(fn, 'POS ', pt) / 1e3
            if do_RSD:
                vel = RG.read_field(fn, 'VEL ', pt)
                pos = oracle.pos_redshift_space(pos, vel, box, h.Hubble, h.redshift, axis)
            W = None
            if Masses[pt] != 0:
                if not single:
                    W = np.ones(hh.npart[pt], np.float32) * np.float32(Masses[pt])
            else:
                W = (RG.read_field(fn, 'MASS', pt) * 1e10).astype(np.float32)
            oracle.ma_f32(pos, grid, box, MAS, W)
            allpos.append(pos)
    return grid, np.concatenate(allpos), box


@pytest.mark.parametrize('do_RSD', [False, True])
def test_density_field_gadget(oracle, tmp_path, do_RSD):
    from pylians3_amd import MAS_gadget
    snap = _snapshot(tmp_path)
    dims = 32
    for ptypes in ([1], [1, 2, 4]):
        print('case gadget', ptypes, do_RSD, flush=True)
        got = MAS_gadget.density_field_gadget(snap, ptypes, dims, 'CIC', do_RSD=do_RSD, axis=1,
                                              verbose=False)
        ref, allpos, box = _oracle_density(oracle, snap, ptypes, dims, do_RSD, 1)
        # several deposits (files, types) per grid: the reference adds each file's
        # float32 contributions in turn, the GPU rounds each deposit once
        assert_cells(got, ref, what='gadget %s' % ptypes,
                     nadds=contrib_counts(allpos, dims, box, 'CIC') + 8)


def test_pk_comp_and_pk_gadget_files(oracle, tmp_path):
    from pylians3_amd import Pk_snapshot as PS
    from pylians3_amd import readgadget as RG
    snap = _snapshot(tmp_path, nfiles=2, fmt=1, redshift=1.0, seed=3)
    dims, axis = 32, 0
    out = str(tmp_path)
    r = PS.Pk_comp(snap, 1, dims, True, axis, 1, out)
    fn = out + '/Pk_CDM_RS_axis=0_z=1.000.dat'
    data = np.loadtxt(fn)
    assert data.shape == (len(r.k3D), 5)
    # oracle: read all of type 1, shift, CIC, delta, Pk
    h = RG.header(snap)
    box = h.boxsize / 1e3
    pos = RG.read_block(snap, 'POS ', [1]) / 1e3
    vel = RG.read_block(snap, 'VEL ', [1])
    Hub = 100.0 * np.sqrt(h.omega_m * (1.0 + h.redshift) ** 3 + h.omega_l)
    pos = oracle.pos_redshift_space(pos, vel, box, Hub, h.redshift, axis)
    g = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, g, box, 'CIC')
    mean = np.float32(len(pos) / dims ** 3)
    delta = g / mean - np.float32(1.0)
    ref = oracle.pk(delta, box, axis=axis, MAS='CIC')
    np.testing.assert_array_equal(data[:, 4], ref.Nmodes3D)
    np.testing.assert_allclose(data[:, 0], ref.k3D, rtol=1e-7)
    P0 = ref.Pk[:, 0]
    assert_bins(data[:, 1], P0, P0, what='Pk_comp l=0')
    assert_bins(data[:, 2], ref.Pk[:, 1], P0, what='Pk_comp l=2')
    # multi-type: auto, cross and total files appear
    PS.Pk_Gadget(snap, dims, [1, 2], False, axis, 1, out)
    for name in ('Pk_CDM_z=1.000.dat', 'Pk_NU_z=1.000.dat', 'Pk_CDMNU_z=1.000.dat',
                 'Pk_CDM+NU_z=1.000.dat'):
        assert os.path.exists(os.path.join(out, name)), name
    cdm = np.loadtxt(os.path.join(out, 'Pk_CDM_z=1.000.dat'))
    g = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(RG.read_block(snap, 'POS ', [1]) / 1e3, g, box, 'CIC')
    ref = oracle.pk(g / mean - np.float32(1.0), box, axis=axis, MAS='CIC')
    assert_bins(cdm[:, 1], ref.Pk[:, 0], ref.Pk[:, 0], what='Pk_Gadget CDM')
