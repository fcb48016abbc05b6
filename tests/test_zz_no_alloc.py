"""lbm_solver_mass_energy_noalloc through the C-ABI: the allocation-free
variant of lbm_solver_mass_energy gives the same FP64 mass and kinetic
energy, bit for bit, on consecutive calls whose second input is smaller,
and rejects a solver beyond its fixed scratch with LBM_ERR_CAPACITY.

Stated tolerance: 0 (same kernels and summation order as the allocating
entry point)."""
import ctypes as C
import sys

import pytest

pytestmark = pytest.mark.gpu

from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from turbulent_lbm_multigpu_amd.solver import LbmSolver  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(lib, s):
    m0, k0, m1, k1 = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    assert lib.lbm_solver_mass_energy(s._h, C.byref(m0), C.byref(k0)) == 0
    assert lib.lbm_solver_mass_energy_noalloc(s._h, C.byref(m1), C.byref(k1)) == 0, \
        lib.lbm_solver_error(s._h).decode()
    return (m0.value, k0.value), (m1.value, k1.value)


@pytest.mark.parametrize("storage", [_capi.STORAGE_F32, _capi.STORAGE_F16_SHIFTED, _capi.STORAGE_F64])
def test_noalloc_equals_allocating_entry_point(storage):
    lib = _capi.load_library()
    sizes = [(130, 97, 66), (23, 19, 17)]   # second call on a smaller solver
    for dims in sizes:
        say(f"case: mass/energy noalloc storage {storage} dims {dims}")
        s = LbmSolver(dims, (1,) * 6, inv_tau=1 / 0.56, tau=0.56, smagorinsky_cs=0.17, lid_velocity=0.07,
                      storage=storage)
        s.init_perturbed(1e-3, 77)
        s.run(5)
        s.wait()
        want, got = _both(lib, s)
        assert got == want, (got, want)
        assert want[0] > 0.5 * (dims[0] - 2) * (dims[1] - 2) * (dims[2] - 2) and want[1] > 0
        # the next call sees the previous call's partials in the scratch: still the same
        s.run(3)
        s.wait()
        want, got = _both(lib, s)
        assert got == want, (got, want)
        s.close()


def test_noalloc_rejects_beyond_its_scratch():
    lib = _capi.load_library()
    dims = (324, 324, 324)   # 34,012,224 cells > 2^25
    say(f"case: mass/energy noalloc rejects {dims}")
    s = LbmSolver(dims, (1,) * 6, inv_tau=1 / 0.6, tau=0.6, storage=_capi.STORAGE_F16_SHIFTED)
    m, k = C.c_double(), C.c_double()
    rc = lib.lbm_solver_mass_energy_noalloc(s._h, C.byref(m), C.byref(k))
    assert rc == _capi.LBM_ERR_CAPACITY
    assert "scratch" in lib.lbm_solver_error(s._h).decode()
    # the allocating entry point still serves it
    assert lib.lbm_solver_mass_energy(s._h, C.byref(m), C.byref(k)) == 0 and m.value > 0
    s.close()
