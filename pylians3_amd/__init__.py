"""pylians3_amd: MI355X-native density-field (MA) and power-spectrum (Pk/XPk) path.

    from pylians3_amd import MAS_library as MASL
    from pylians3_amd import Pk_library as PKL

The compute runs in libpylians_hip.so (HIP kernels for gfx950 + rocFFT), built
in-tree by ``python -m pylians3_amd.build`` and loaded through ctypes.
"""
from . import _native  # noqa: F401

__version__ = '0.1.0'
