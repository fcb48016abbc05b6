"""MI355X-native D3Q19 lattice-Boltzmann hot path (gfx950 HIP kernels behind a C-ABI).

The compute path is liblbm_hip.so (built by ``build.py``); this package holds
the Python host layer over its C-ABI: ``LbmSolver`` / ``LbmGroup`` handles,
the reference-shaped ``CLbmSolver`` / ``CController`` / ``CManager`` classes,
the parametrization and the configuration surface.
"""
from . import _capi  # noqa: F401
from ._capi import LbmError, load_library  # noqa: F401

__all__ = ["LbmError", "load_library"]
