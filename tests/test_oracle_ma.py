"""Pins the mass-assignment oracle (oracle/ma_oracle.c) before it is trusted.

Checks chosen so that a dropped term, a wrong node offset, a wrong wrap or a
changed rounding fails one of them: closed-form single-particle footprints,
partition of unity, an independent float64 brute force built from the
B-spline definitions, periodic wrap, and bitwise equality with the reference's
own MAS_c.c compiled into oracle/_ref (library/MAS_library/MAS_c.c:8-266).
"""
import numpy as np
import pytest

SCHEMES = ['NGP', 'CIC', 'TSC', 'PCS']


def bspline(r, mas):
    """Mass-assignment window of periodic distance r (cell units), float64."""
    a = np.abs(r)
    if mas == 'CIC':
        return np.maximum(0.0, 1.0 - a)
    if mas == 'TSC':
        return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    if mas == 'PCS':
        return np.where(a < 1.0, (4.0 - 6.0 * a * a + 3.0 * a ** 3) / 6.0,
                        np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))
    raise ValueError(mas)


def brute_grid(pos, dims, box, mas, W=None):
    """Independent float64 grid: sum over particles of W * prod_axes window(node - s)."""
    pos = pos.astype(np.float64)
    n, coord = pos.shape
    s = pos * (dims / box)
    nodes = np.arange(dims, dtype=np.float64)
    grid = np.zeros((dims,) * coord)
    w = np.ones(n) if W is None else W.astype(np.float64)
    for p in range(n):
        per_axis = []
        for ax in range(coord):
            if mas == 'NGP':
                v = np.zeros(dims)
                v[int(np.floor(s[p, ax] + 0.5)) % dims] = 1.0
            else:
                r = nodes - s[p, ax]
                r = r - dims * np.round(r / dims)          # nearest periodic image
                v = bspline(r, mas)
            per_axis.append(v)
        if coord == 3:
            grid += w[p] * np.einsum('i,j,k->ijk', *per_axis)
        else:
            grid += w[p] * np.outer(*per_axis)
    return grid


@pytest.mark.parametrize('mas,x,nodes,weights', [
    ('CIC', 5.25, [5, 6], [0.75, 0.25]),
    ('CIC', 0.5, [0, 1], [0.5, 0.5]),
    ('TSC', 5.25, [4, 5, 6], [0.03125, 0.6875, 0.28125]),
    ('TSC', 5.0, [4, 5, 6], [0.125, 0.75, 0.125]),
    ('PCS', 5.5, [4, 5, 6, 7], [1 / 48, 23 / 48, 23 / 48, 1 / 48]),
    ('PCS', 5.0, [4, 5, 6, 7], [1 / 6, 4 / 6, 1 / 6, 0.0]),
    ('NGP', 5.49, [5], [1.0]),
    ('NGP', 5.5, [6], [1.0]),
])
def test_single_axis_closed_form(oracle, mas, x, nodes, weights):
    idx, C = oracle.axis_weights(x, 1.0, 16, mas)
    assert idx == nodes
    np.testing.assert_allclose(C, weights, rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize('mas', SCHEMES)
def test_partition_of_unity(oracle, mas):
    rng = np.random.default_rng(7)
    for x in rng.random(2000) * 64:
        _, C = oracle.axis_weights(float(np.float32(x)), 1.0, 64, mas)
        assert abs(C.astype(np.float64).sum() - 1.0) < 3e-7


@pytest.mark.parametrize('mas', SCHEMES)
@pytest.mark.parametrize('coord', [2, 3])
@pytest.mark.parametrize('weighted', [False, True])
def test_brute_force_tiny_grid(oracle, mas, coord, weighted):
    rng = np.random.default_rng(11)
    dims, box = 6, 3.0
    pos = (rng.random((300, coord)) * box).astype(np.float32)
    W = (rng.random(300) + 0.5).astype(np.float32) if weighted else None
    ref = brute_grid(pos, dims, box, mas, W)
    g = np.zeros((dims,) * coord, np.float32)
    oracle.ma_f32(pos, g, box, mas, W, cython2d=True, renormalize_2D=True)
    np.testing.assert_allclose(g, ref, rtol=2e-6, atol=2e-6 * ref.mean())
    g64 = oracle.ma_f64(pos, dims, box, mas, W)
    np.testing.assert_allclose(g64, ref, rtol=1e-12, atol=1e-12 * ref.mean())


@pytest.mark.parametrize('mas', SCHEMES)
def test_mass_conservation(oracle, mas):
    # library/tests/test.py:16-33 criterion: |sum/elements - 1| < 1e-5
    rng = np.random.default_rng(3)
    dims = 32
    pos = rng.random((dims ** 3, 3)).astype(np.float32)
    g = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(pos, g, 1.0, mas)
    assert abs(np.sum(g, dtype=np.float64) / dims ** 3 - 1.0) < 1e-5
    pos2 = rng.random((dims ** 2, 2)).astype(np.float32)
    g2 = np.zeros((dims,) * 2, np.float32)
    oracle.ma_f32(pos2, g2, 1.0, mas)
    assert abs(np.sum(g2, dtype=np.float64) / dims ** 2 - 1.0) < 1e-5


@pytest.mark.parametrize('mas', SCHEMES)
def test_periodic_wrap(oracle, mas):
    dims, box = 8, 8.0
    # x == BoxSize behaves like x == 0
    a = np.zeros((dims,) * 3, np.float32)
    b = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(np.array([[box, 3.0, 4.0]], np.float32), a, box, mas)
    oracle.ma_f32(np.array([[0.0, 3.0, 4.0]], np.float32), b, box, mas)
    np.testing.assert_array_equal(a, b)
    # a particle just below the box edge deposits on both ends of the axis
    c = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(np.array([[box - 0.25, 3.0, 4.0]], np.float32), c, box, mas)
    plane = c.sum(axis=(1, 2))
    if mas != 'NGP':
        assert plane[dims - 1] > 0 and plane[0] > 0
    assert abs(plane.sum() - 1.0) < 1e-6
    # translation by whole cells shifts the grid
    p = np.array([[2.3, 4.6, 1.7]], np.float32)
    d0 = np.zeros((dims,) * 3, np.float32)
    d1 = np.zeros((dims,) * 3, np.float32)
    oracle.ma_f32(p, d0, box, mas)
    oracle.ma_f32(p + np.float32(3.0), d1, box, mas)
    np.testing.assert_allclose(np.roll(d0, 3, axis=(0, 1, 2)), d1, atol=1e-6)


def test_inv_cell_size_float_equals_double_rounded():
    # Cython computes dims/BoxSize in float (MAS_library.pyx:135); MAS_c.c:19 in
    # double then stores to float.  Both round to the same float.
    rng = np.random.default_rng(5)
    for dims in (2, 3, 7, 64, 100, 128, 255, 256, 512, 1000, 1024):
        for box in list(rng.random(300) * 5000) + [1.0, 500.0, 1000.0, 2000.0, 3.0]:
            b = np.float32(box)
            f = np.float32(np.float32(dims) / b)
            d = np.float32(np.float64(dims) / np.float64(b))
            assert f == d


@pytest.mark.parametrize('mas', SCHEMES)
@pytest.mark.parametrize('weighted', [False, True])
def test_bitwise_vs_reference_mas_c(oracle, mas, weighted):
    if not oracle.ref_available('strict'):
        pytest.skip('oracle/_ref not built (reference not mounted)')
    rng = np.random.default_rng(21)
    for coord, dims, n in ((3, 24, 20000), (2, 40, 20000)):
        pos = (rng.random((n, coord)) * 250.0).astype(np.float32)
        pos[:5] = np.float32(250.0)          # edge: x == BoxSize
        pos[5:10] = 0.0
        W = (rng.random(n) * 3).astype(np.float32) if weighted else None
        a = np.zeros((dims,) * coord, np.float32)
        b = np.zeros((dims,) * coord, np.float32)
        oracle.ma_f32(pos, a, 250.0, mas, W, cython2d=False)
        oracle.ma_reference(pos, b, 250.0, mas, W, flavour='strict')
        np.testing.assert_array_equal(a, b)
        if oracle.ref_available('shipped'):
            c = np.zeros((dims,) * coord, np.float32)
            oracle.ma_reference(pos, c, 250.0, mas, W, flavour='shipped')
            np.testing.assert_allclose(c, a, rtol=1e-5, atol=1e-5 * np.abs(a).mean())


@pytest.mark.parametrize('mas', ['CIC', 'TSC', 'PCS'])
def test_cython_2d_path_is_F_times_native(oracle, mas):
    # MA() on 2D runs the 3D kernel on a (d,d,1) view: F repeats, then /F
    rng = np.random.default_rng(9)
    pos = (rng.random((5000, 2)) * 10).astype(np.float32)
    native = np.zeros((16, 16), np.float32)
    oracle.ma_f32(pos, native, 10.0, mas, cython2d=False)
    rep = np.zeros((16, 16), np.float32)
    oracle.ma_f32(pos, rep, 10.0, mas, cython2d=True, renormalize_2D=False)
    F = {'CIC': 2, 'TSC': 3, 'PCS': 4}[mas]
    np.testing.assert_allclose(rep, F * native, rtol=1e-5)
    ren = np.zeros((16, 16), np.float32)
    oracle.ma_f32(pos, ren, 10.0, mas, cython2d=True, renormalize_2D=True)
    np.testing.assert_allclose(ren, native, rtol=1e-5)


def test_accumulates_into_existing_grid(oracle):
    rng = np.random.default_rng(2)
    pos = (rng.random((1000, 3)) * 5).astype(np.float32)
    g = np.full((8, 8, 8), 2.0, np.float32)
    oracle.ma_f32(pos, g, 5.0, 'CIC')
    h = np.zeros((8, 8, 8), np.float32)
    oracle.ma_f32(pos, h, 5.0, 'CIC')
    np.testing.assert_allclose(g, h + 2.0, rtol=1e-6)


def test_redshift_space_oracle(oracle):
    # redshift_space_library.pyx:29-46: s = r + v (1+z)/H, then wrap
    pos = np.array([[10.0, 20.0, 30.0], [1.0, 2.0, 999.0], [5.0, 5.0, 5.0]], np.float32)
    vel = np.array([[100.0, 0.0, 0.0], [-300.0, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
    out = oracle.pos_redshift_space(pos.copy(), vel, 1000.0, 100.0, 0.0, 0)
    assert out[0, 0] == np.float32(11.0)
    assert out[1, 0] == np.float32(np.float32(1.0 - 3.0) + np.float32(1000.0))
    np.testing.assert_array_equal(out[:, 1:], pos[:, 1:])
    out2 = oracle.pos_redshift_space(np.array([[999.5, 0, 0]], np.float32),
                                     np.array([[100.0, 0, 0]], np.float32), 1000.0, 100.0, 0.0, 0)
    assert out2[0, 0] == np.float32(np.fmod(np.float32(1000.5), np.float32(1000.0)))


def test_cic_interp_oracle_adjoint_and_closed_form(oracle):
    # CIC_interp is the adjoint of CIC: <MA_CIC(p, w), g> = sum_i w_i interp(g, p_i)
    rng = np.random.default_rng(31)
    dims, box = 12, 30.0
    pos = (rng.random((500, 3)) * box).astype(np.float32)
    w = rng.random(500)
    g = rng.standard_normal((dims,) * 3).astype(np.float32)
    grid64 = oracle.ma_f64(pos, dims, box, 'CIC', w.astype(np.float32))
    lhs = np.sum(grid64 * g.astype(np.float64))
    rhs = np.sum(w.astype(np.float32).astype(np.float64) * oracle.cic_interp(g, box, pos))
    assert abs(lhs - rhs) < 1e-5 * np.sum(np.abs(grid64 * g))
    # a linear field is reproduced exactly between nodes (x = 2.5 cells -> 2.5 along x)
    lin = np.tile(np.arange(dims, dtype=np.float32)[:, None, None], (1, dims, dims))
    v = oracle.cic_interp(lin, box, np.array([[2.5 * box / dims, 1.0, 1.0]], np.float32))
    assert abs(v[0] - 2.5) < 1e-5


@pytest.mark.parametrize('mas', ['NGP', 'CIC', 'TSC', 'PCS'])
@pytest.mark.parametrize('coord', [2, 3])
def test_threaded_f64_equals_serial(oracle, mas, coord):
    """The slab-phased threaded FP64 MA equals the serial one to double rounding
    and is deterministic, including grids too thin for two slabs per thread."""
    rng = np.random.default_rng(21)
    for dims in (3, 8, 31, 64):
        pos = (rng.random((150000, coord)) * 77.0).astype(np.float32)
        W = rng.lognormal(0, 1, 150000).astype(np.float32)
        a = oracle.ma_f64(pos, dims, 77.0, mas, W)
        b = oracle.ma_f64(pos, dims, 77.0, mas, W, threads=6)
        c = oracle.ma_f64(pos, dims, 77.0, mas, W, threads=6)
        np.testing.assert_array_equal(b, c)
        np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-13 * np.abs(a).max())


@pytest.mark.skipif(not __import__('oracle.oracle', fromlist=['x']).cython_available(),
                    reason='oracle/_ref/cy_strict not built (needs the reference sources: make -C oracle cython)')
@pytest.mark.parametrize('mas', ['NGP', 'CIC', 'TSC', 'PCS'])
@pytest.mark.parametrize('coord', [2, 3])
@pytest.mark.parametrize('weighted', [False, True])
@pytest.mark.parametrize('renorm', [True, False])
def test_oracle_equals_reference_cython_ma(oracle, mas, coord, weighted, renorm):
    """The float32 restatement is bitwise the reference's own MA() (MAS_library.pyx
    translated by Cython and compiled with strict IEEE flags into oracle/_ref): the
    serial loops MA() runs, the 2D expand_dims path and its renormalisation of a
    pre-filled grid (MAS_library.pyx:57-545)."""
    rng = np.random.default_rng(7 + coord)
    box, dims, n = 90.0, 20, 30000
    pos = (rng.random((n, coord)) * box).astype(np.float32)
    pos[:5] = [[0.0] * coord, [box] * coord, [box * (1 - 1e-7)] * coord,
               [box / dims * 3.5] * coord, [box / dims * 7.0] * coord]
    W = rng.lognormal(0, 1, n).astype(np.float32) if weighted else None
    init = rng.random((dims,) * coord).astype(np.float32)
    a = init.copy()
    oracle.ma_f32(pos, a, box, mas, W, renormalize_2D=renorm)
    b = init.copy()
    oracle.ma_cython(pos, b, box, mas, W, renormalize_2D=renorm)
    np.testing.assert_array_equal(a, b)
