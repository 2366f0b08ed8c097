"""Write the golden fixtures tests/golden/{ma,pk}_golden.npz.

Inputs are seeded; expected outputs come from the oracle, which is pinned by
tests/test_oracle_*.py.  When oracle/_ref is built (the reference's own
MAS_c.c compiled from /root/reference by oracle/Makefile), the 3D and native-2D
grids are also checked here to be bitwise identical to the reference's output
before they are written, so the MA fixtures are the reference's own results.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SCHEMES = ['NGP', 'CIC', 'TSC', 'PCS']


def golden_positions():
    rng = np.random.default_rng(100)
    box = 100.0
    pos = (rng.random((3000, 3)) * box).astype(np.float32)
    edge = np.array([0.0, box, np.nextafter(np.float32(box), np.float32(0)), 6.25, 50.0],
                    np.float32)
    pos[:5] = edge[:, None]
    pos[5:10, 0] = edge
    W = rng.lognormal(0.0, 1.0, 3000).astype(np.float32)
    pos2 = (np.random.default_rng(102).random((2000, 2)) * box).astype(np.float32)
    return box, pos, W, pos2


def main():
    box, pos, W, pos2 = golden_positions()
    ma = {'box': np.float64(box), 'pos': pos, 'W': W, 'pos2': pos2}
    for mas in SCHEMES:
        for tag, w in (('u', None), ('w', W)):
            g = np.zeros((16,) * 3, np.float32)
            O.ma_f32(pos, g, box, mas, w)
            if O.ref_available('strict'):
                r = np.zeros_like(g)
                O.ma_reference(pos, r, box, mas, w, flavour='strict')
                assert np.array_equal(g, r), mas
            ma['grid3d_%s_%s' % (mas, tag)] = g
            g2 = np.zeros((24, 24), np.float32)
            O.ma_f32(pos2, g2, box, mas, None if w is None else w[:2000], cython2d=True,
                     renormalize_2D=True)
            ma['grid2d_%s_%s' % (mas, tag)] = g2
            n2 = np.zeros((24, 24), np.float32)
            O.ma_f32(pos2, n2, box, mas, None if w is None else w[:2000], cython2d=False)
            if O.ref_available('strict'):
                r2 = np.zeros_like(n2)
                O.ma_reference(pos2, r2, box, mas, None if w is None else w[:2000], flavour='strict')
                assert np.array_equal(n2, r2), mas
            ma['grid2dnative_%s_%s' % (mas, tag)] = n2
    np.savez_compressed(os.path.join(HERE, 'ma_golden.npz'), **ma)

    g = ma['grid3d_CIC_u']
    delta = g / np.float32(g.mean()) - np.float32(1.0)
    gw = ma['grid3d_TSC_w']
    delta2 = gw / np.float32(gw.mean()) - np.float32(1.0)
    pk = {'box': np.float64(box), 'delta': delta, 'delta2': delta2}
    for axis in (0, 1, 2):
        r = O.pk(delta, box, axis=axis, MAS='CIC')
        for name in ('k3D', 'Pk', 'Nmodes3D', 'k1D', 'Pk1D', 'Nmodes1D', 'kpar', 'kper', 'Pk2D',
                     'Nmodes2D', 'Pkphase'):
            pk['pk_ax%d_%s' % (axis, name)] = getattr(r, name)
    x = O.xpk([delta, delta2], box, axis=2, MAS=['CIC', 'TSC'])
    for name in ('Pk', 'XPk', 'Pk1D', 'PkX1D', 'Pk2D', 'PkX2D', 'Nmodes3D'):
        pk['xpk_%s' % name] = getattr(x, name)
    np.savez_compressed(os.path.join(HERE, 'pk_golden.npz'), **pk)
    print('wrote', sorted(ma)[:3], '...', len(ma), 'MA arrays and', len(pk), 'Pk arrays')


if __name__ == '__main__':
    main()
