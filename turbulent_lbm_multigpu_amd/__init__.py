"""MI355X-native D3Q19 lattice-Boltzmann hot path (gfx950 HIP kernels behind a C-ABI).

The compute path is liblbm_hip.so (built by ``build.py``).  The
reference-shaped host classes (``CLbmSolver``, ``CLbmSkeleton``, ``CController``,
``CManager``, ``CConfiguration``) are C++ headers in ``csrc/host`` over the
same C-ABI, driven by the ``lbm_main`` command-line program.  This package is
the Python layer over the C-ABI: ``solver`` (``LbmSolver`` / ``LbmGroup``
handles), ``decomposition``, ``batch``, ``serving``, ``precision`` and
``profiling``.
"""
from . import _capi  # noqa: F401
from ._capi import LbmError, load_library  # noqa: F401

__all__ = ["LbmError", "load_library"]
