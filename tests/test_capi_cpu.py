"""CPU-side checks of the C-ABI library: it loads without a GPU, exports every
symbol include/lbm_capi.h declares, and rejects bad arguments before any
device call."""
import ctypes as C
import os
import re

import pytest

from turbulent_lbm_multigpu_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lbm_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(lbm_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _capi.load_library()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/lbm_capi.h but not exported"
    assert set(_capi.EXPORTED_SYMBOLS) <= set(names)


def test_version_and_defaults():
    lib = _capi.load_library()
    assert b"gfx950" in lib.lbm_version()
    cfg = _capi.SolverConfig()
    lib.lbm_solver_config_default(C.byref(cfg))
    assert cfg.bc[0] == _capi.FLAG_OBSTACLE and cfg.work_group_size == 256
    assert abs(cfg.tau - 0.953575) < 1e-12   # src/CLbmSkeleton.hpp:142


@pytest.mark.parametrize("field,value,code", [
    ("dims", (2, 8, 8), _capi.LBM_ERR_INVALID_ARGUMENT),
    ("bc", (1, 1, 1, 1, 1, 300), _capi.LBM_ERR_INVALID_ARGUMENT),
    ("storage", 9, _capi.LBM_ERR_INVALID_ARGUMENT),
    ("work_group_size", 100, _capi.LBM_ERR_INVALID_ARGUMENT),
    ("kernel_variant", 10, _capi.LBM_ERR_INVALID_ARGUMENT),   # variants: 0-3, 8, 11, 16 (each | 4)
    ("kernel_variant", 13, _capi.LBM_ERR_INVALID_ARGUMENT),   # 9 | 4
    ("kernel_variant", 9, _capi.LBM_ERR_INVALID_ARGUMENT),
    ("kernel_variant", 24, _capi.LBM_ERR_INVALID_ARGUMENT),
    ("kernel_variant", -1, _capi.LBM_ERR_INVALID_ARGUMENT),
    ("inv_tau", -1.0, _capi.LBM_ERR_PARAMETRIZATION),
])
def test_create_rejects_bad_config_without_device(field, value, code):
    lib = _capi.load_library()
    cfg = _capi.SolverConfig()
    lib.lbm_solver_config_default(C.byref(cfg))
    if isinstance(value, tuple):
        arr = getattr(cfg, field)
        for i, v in enumerate(value):
            arr[i] = v
    else:
        setattr(cfg, field, value)
    h = C.c_void_p()
    rc = lib.lbm_solver_create(C.byref(cfg), C.byref(h))
    assert rc == code
    assert h.value is None
    assert len(lib.lbm_last_error()) > 0


def test_group_rejects_empty():
    lib = _capi.load_library()
    h = C.c_void_p()
    assert lib.lbm_group_create(None, 0, C.byref(h)) == _capi.LBM_ERR_INVALID_ARGUMENT
