"""Host-side behaviour of the drop-in modules that runs before any GPU work:
the reference's SystemExit / ValueError paths (MAS_library.pyx:61-66,81-82;
Pk_library.pyx:553-560 and Cython buffer checks), helper functions and the
normalisation constants."""
import numpy as np
import pytest

from pylians3_amd import MAS_library as MASL
from pylians3_amd import Pk_library as PKL


def test_ma_dimension_mismatch_exits(capsys):
    pos = np.zeros((10, 3), np.float32)
    with pytest.raises(SystemExit):
        MASL.MA(pos, np.zeros((8, 8), np.float32), 1.0, 'CIC')
    assert 'pos have 3 dimensions and the density 2!!!' in capsys.readouterr().out


def test_ma_invalid_scheme_exits(capsys):
    with pytest.raises(SystemExit):
        MASL.MA(np.zeros((10, 3), np.float32), np.zeros((8, 8, 8), np.float32), 1.0, 'XYZ')
    assert 'option not valid!!!' in capsys.readouterr().out


def test_ma_dtype_errors():
    with pytest.raises(ValueError, match='float32_t'):
        MASL.MA(np.zeros((10, 3)), np.zeros((8, 8, 8), np.float32), 1.0, 'CIC')
    with pytest.raises(ValueError, match='float32_t'):
        MASL.MA(np.zeros((10, 3), np.float32), np.zeros((8, 8, 8)), 1.0, 'CIC')
    with pytest.raises(ValueError, match='float32_t'):
        MASL.MA(np.zeros((10, 3), np.float32), np.zeros((8, 8, 8), np.float32), 1.0, 'CIC',
                W=np.ones(10))


def test_ma_unsupported_coord_is_noop():
    g = np.zeros(8, np.float32)
    assert MASL.MA(np.zeros((4, 1), np.float32), g, 1.0, 'CIC') is None
    assert not g.any()


def test_c_wrappers_require_contiguity():
    pos = np.zeros((10, 6), np.float32)[:, ::2]
    with pytest.raises(ValueError, match='C-contiguous'):
        MASL.CICc3D(pos, np.zeros((8, 8, 8), np.float32), 1.0, 1)


def test_float_type():
    assert MASL.FLOAT_type() is np.float32


def test_pk_helpers_match_reference_formulas():
    kF, kN, kpar, kper, kmax = PKL.frequencies(1000.0, 512)
    assert kF == 2.0 * np.pi / 1000.0 and kN == 256 * kF
    assert (kpar, kper, kmax) == (256, 362, 443)
    assert PKL.frequencies_2D(10.0, 64)[2:] == (32, 32, 45)
    assert [PKL.MAS_function(m) for m in ('NGP', 'CIC', 'TSC', 'PCS', 'None')] == [1, 2, 3, 4, 0]
    assert PKL.MAS_correction(0.0, 4) == 1.0
    assert PKL.MAS_correction(0.5, 2) == pytest.approx((0.5 / np.sin(0.5)) ** 2, rel=1e-15)


def test_check_number_modes(capsys):
    PKL.check_number_modes(np.array([(16 ** 3 - 8) // 2 + 8.0]), 16)
    with pytest.raises(SystemExit):
        PKL.check_number_modes(np.array([5.0]), 16)
    assert 'Not all modes counted' in capsys.readouterr().out
    PKL.check_number_modes_2D(np.array([(15 ** 2 - 1) // 2 + 1.0]), 15)


def test_pk_field_checks():
    with pytest.raises(ValueError, match='float32_t'):
        PKL.Pk(np.zeros((8, 8, 8)), 1.0, verbose=False)
    with pytest.raises(ValueError, match='dimensions'):
        PKL.Pk(np.zeros((8, 8), np.float32), 1.0, verbose=False)


def test_xpk_size_mismatch_exits(capsys):
    with pytest.raises(SystemExit):
        PKL.XPk([np.zeros((8, 8, 8), np.float32), np.zeros((4, 4, 4), np.float32)], 1.0,
                MAS=['CIC', 'CIC'])
    assert 'Fields have different grid sizes!!!' in capsys.readouterr().out
    with pytest.raises(TypeError):
        PKL.XPk([np.zeros((8, 8, 8), np.float32)] * 2, 1.0, MAS=None)
