"""CPU-side checks of the batching and serving host logic: the seeded
workloads follow their stated distributions, and the C-ABI sizes cases and
solver arenas without a device."""
import ctypes as C
import math

import numpy as np

from turbulent_lbm_multigpu_amd import _capi
from turbulent_lbm_multigpu_amd import batch
from turbulent_lbm_multigpu_amd import serving


def test_cavity_set_distribution():
    cases = batch.make_cavities(512, seed=20260101)
    assert len(cases) == 512
    dims = np.array([c.dims for c in cases])
    assert dims.min() == 16 and dims.max() == 96
    # uniform edges: mean 56, mean cells ~ 56^3
    assert abs(dims.mean() - 56) < 1.5
    assert 0.9 < np.prod(dims, axis=1).mean() / 56 ** 3 < 1.1
    lids = np.array([c.lid for c in cases])
    assert lids.min() >= 0.02 and lids.max() <= 0.1
    taus = np.array([c.tau for c in cases])
    assert taus.min() >= 0.5002
    assert abs(np.median(taus) - 0.52) < 0.002
    assert batch.make_cavities(4, seed=3) == batch.make_cavities(4, seed=3)


def test_request_stream_distribution():
    reqs = serving.request_stream(4000, rate=5.0, seed=7)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 0.2) < 0.02                     # Poisson, mean gap 1/rate
    edges = np.array([r.edge for r in reqs])
    assert edges.min() >= 32 and edges.max() <= 192
    # log-uniform: the median edge is the geometric mean of the bounds
    assert abs(np.median(edges) - math.sqrt(32 * 192)) < 6
    steps = np.array([r.steps for r in reqs])
    assert steps.min() >= 500 and steps.max() <= 5000 and abs(steps.mean() - 2750) < 60
    kinds = {r.kind for r in reqs}
    assert kinds == {"cavity", "channel"}
    assert {r.storage for r in reqs} == {"f32", "f16"}
    at_once = serving.request_stream(10, rate=0.0, seed=7)
    assert all(r.arrival == 0.0 for r in at_once)


def test_case_bytes_without_device():
    c32 = serving.make_case("cavity", (64, 64, 64), 10, "f32")
    c16 = serving.make_case("channel", (64, 64, 64), 10, "f16")
    b32, b16 = serving.case_bytes(c32), serving.case_bytes(c16)
    cells = 64 ** 3
    # 19 populations + 1 flag byte per cell, plus tables and staging
    assert 77 * cells < b32 < 100 * cells
    assert 39 * cells < b16 < 60 * cells
    lib = _capi.load_library()
    bad = serving.make_case("cavity", (64, 64, 64), 10, "f32")
    bad.kind = 7
    n = C.c_size_t()
    assert lib.lbm_server_case_bytes(C.byref(bad), C.byref(n)) == _capi.LBM_ERR_INVALID_ARGUMENT


def test_required_bytes_grows_with_storage_and_fields():
    lib = _capi.load_library()
    cfg = _capi.SolverConfig()
    lib.lbm_solver_config_default(C.byref(cfg))
    for a in range(3):
        cfg.dims[a] = 40
    sizes = {}
    for st in (_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_F32, _capi.STORAGE_F64):
        cfg.storage = st
        n = C.c_size_t()
        assert lib.lbm_solver_required_bytes(C.byref(cfg), C.byref(n)) == 0
        sizes[st] = n.value
        assert n.value % 256 == 0
    assert sizes[_capi.STORAGE_F16_SHIFTED] < sizes[_capi.STORAGE_F32] < sizes[_capi.STORAGE_F64]
    cfg.storage = _capi.STORAGE_F32
    cfg.store_velocity = 1
    n = C.c_size_t()
    lib.lbm_solver_required_bytes(C.byref(cfg), C.byref(n))
    assert n.value >= sizes[_capi.STORAGE_F32] + 3 * 4 * 40 ** 3
