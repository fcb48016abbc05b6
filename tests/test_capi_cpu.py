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
    ("kernel_variant", 10, _capi.LBM_ERR_INVALID_ARGUMENT),   # variants: 0-3, 8, 11, 16 (each | 4 or | 32)
    ("kernel_variant", 13, _capi.LBM_ERR_INVALID_ARGUMENT),   # 9 | 4
    ("kernel_variant", 39, _capi.LBM_ERR_INVALID_ARGUMENT),   # 3 | 4 | 32: no map and chunks
    ("kernel_variant", 64, _capi.LBM_ERR_INVALID_ARGUMENT),
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


def _variant_accepted_by(lib, variant, wg=256):
    """(solver, group, server) acceptance of one launch shape, decided before
    any device call: a shape that passes the checks fails later at the device
    (no GPU here) with a different code."""
    cfg = _capi.SolverConfig()
    lib.lbm_solver_config_default(C.byref(cfg))
    cfg.dims[0] = cfg.dims[1] = cfg.dims[2] = 16
    cfg.kernel_variant = variant
    cfg.work_group_size = wg
    nbytes = C.c_size_t()
    solver_ok = lib.lbm_solver_required_bytes(C.byref(cfg), C.byref(nbytes)) == _capi.LBM_OK
    g = C.c_void_p()
    group_ok = lib.lbm_group_create_empty(0, _capi.STORAGE_F32, wg, variant, C.byref(g)) != \
        _capi.LBM_ERR_INVALID_ARGUMENT
    scfg = _capi.ServerConfig()
    lib.lbm_server_config_default(C.byref(scfg))
    scfg.kernel_variant = variant
    scfg.work_group_size = wg
    scfg.memory_budget = 1 << 24
    srv = C.c_void_p()
    server_ok = lib.lbm_server_create(C.byref(scfg), C.byref(srv)) != _capi.LBM_ERR_INVALID_ARGUMENT
    for h, destroy in ((g, lib.lbm_group_destroy), (srv, lib.lbm_server_destroy)):
        if h.value:
            destroy(h)
    return solver_ok, group_ok, server_ok


def test_solver_group_and_server_accept_the_same_launch_shapes():
    lib = _capi.load_library()
    accepted = []
    for v in range(-1, 66):
        s, g, srv = _variant_accepted_by(lib, v)
        assert s == g == srv, f"variant {v}: solver {s}, group {g}, server {srv}"
        if s:
            accepted.append(v)
    assert accepted == [0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 12, 15, 16, 20, 32, 33, 34, 35, 40, 43, 48]
    for wg in (32, 64, 100, 128, 192, 256, 320, 384, 448, 512, 576, 1024):
        s, g, srv = _variant_accepted_by(lib, 11, wg)
        assert s == g == srv == (wg in (64, 128, 192, 256, 320, 384, 448, 512))


def test_required_bytes_do_not_depend_on_launch_shape():
    """A case sized once (as the server does) fits a solver of any accepted
    work-group size and variant: the tile table is sized for the most tiles."""
    lib = _capi.load_library()
    sizes = set()
    for dims in ((32, 32, 32), (17, 23, 29), (96, 48, 48)):
        sizes.clear()
        for wg in (64, 128, 192, 256, 512):
            for v in (0, 3, 8, 11, 16):
                cfg = _capi.SolverConfig()
                lib.lbm_solver_config_default(C.byref(cfg))
                for a in range(3):
                    cfg.dims[a] = dims[a]
                cfg.work_group_size, cfg.kernel_variant = wg, v
                n = C.c_size_t()
                assert lib.lbm_solver_required_bytes(C.byref(cfg), C.byref(n)) == _capi.LBM_OK
                sizes.add(n.value)
        assert len(sizes) == 1, (dims, sizes)


def test_colocate_rejects_bad_arguments_without_device():
    """lbm_solver_colocate checks its arguments before any device call."""
    import ctypes as C

    lib = _capi.load_library()
    assert lib.lbm_solver_colocate(None, 0) == _capi.LBM_ERR_INVALID_ARGUMENT
    arr = (C.c_void_p * 2)(None, None)
    assert lib.lbm_solver_colocate(arr, 2) == _capi.LBM_ERR_INVALID_ARGUMENT
    assert "null" in lib.lbm_last_error().decode()


def test_alpha_work_group_rejects_null_group():
    """lbm_group_set/get_alpha_work_group check their arguments first."""
    lib = _capi.load_library()
    assert lib.lbm_group_set_alpha_work_group(None, 512) == _capi.LBM_ERR_INVALID_ARGUMENT
    w = C.c_int()
    assert lib.lbm_group_get_alpha_work_group(None, C.byref(w)) == _capi.LBM_ERR_INVALID_ARGUMENT
