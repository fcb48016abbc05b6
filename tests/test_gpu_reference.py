"""Parity against the reference's own kernels.

oracle/_ref holds the reference's OpenCL kernels (lbm_init.cl, lbm_alpha.cl,
lbm_beta.cl, unmodified) compiled for gfx950 by oracle/build_ref.py, and a
launcher replaying the reference host's sequence.  The HIP path must match
them on the same seeded inputs: FP32 rho and u of every FLUID cell within a
relative 1e-5 after N steps (the reference kernels are built by clang's OpenCL
front end, which may contract multiply-adds, so bitwise equality is not
expected).  The CPU oracle is compared to them the same way, which pins the
oracle against the reference's own code.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from turbulent_lbm_multigpu_amd import decomposition as dec  # noqa: E402
from turbulent_lbm_multigpu_amd.solver import LbmSolver  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
RTOL = 1e-5

if not os.path.exists(os.path.join(REF_DIR, "libref_runner.so")):
    pytest.skip("oracle/_ref not built (needs the reference sources at build time)", allow_module_level=True)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def ref_lib():
    lib = C.CDLL(os.path.join(REF_DIR, "libref_runner.so"))
    P = C.c_void_p
    f = C.c_float
    lib.ref_run.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, P, f, f, f, f, f, P, P, C.c_int,
                            P, P, P, P]
    lib.ref_run.restype = C.c_int
    lib.ref_last_error.restype = C.c_char_p
    return lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def run_reference(dims, wg, bc, inv_tau, g, lid, flags, dd0, steps):
    lib = ref_lib()
    nC = int(np.prod(dims))
    dd = np.zeros((19, nC), np.float32)
    vel = np.zeros((3, nC), np.float32)
    dens = np.zeros(nC, np.float32)
    fl = np.zeros(nC, np.int32)
    co = os.path.join(REF_DIR, f"ref_{dims[0]}x{dims[1]}x{dims[2]}_wg{wg}.co").encode()
    rc = lib.ref_run(co, dims[0], dims[1], dims[2], wg, ptr(np.asarray(bc, np.int32)), np.float32(inv_tau),
                     np.float32(g[0]), np.float32(g[1]), np.float32(g[2]), np.float32(lid),
                     ptr(np.ascontiguousarray(flags, np.int32)),
                     ptr(np.ascontiguousarray(dd0, np.float32)) if dd0 is not None else None, steps,
                     ptr(dd), ptr(vel), ptr(dens), ptr(fl))
    assert rc == 0, lib.ref_last_error().decode()
    return dd, vel, dens, fl


def moments(dd, steps):
    """rho, j (float64) from distributions in the parity layout of `steps`."""
    E = orc.E.astype(np.float64)
    f = dd[orc.OPP] if steps % 2 == 0 else dd
    f = f.astype(np.float64)
    return f.sum(0), E.T @ f


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def compare(case, dims, wg, bc, *, tau, lid, g=(0.0, 0.0, 0.0), amp=0.0, seed=0, steps, lid_layer=True):
    say(f"case: {case} dims={dims} wg={wg} steps={steps}")
    inv_tau = float(np.float32(1.0) / np.float32(tau))
    o = orc.OracleDomain(dims, bc, inv_tau=inv_tau, lid_velocity=lid, gravitation=g)
    s = LbmSolver(dims, bc, inv_tau=inv_tau, lid_velocity=lid, gravitation=g, store_velocity=True,
                  store_density=True)
    if amp:
        o.reset(amplitude=amp, seed=seed)
        s.init_perturbed(amp, seed)
    if lid_layer:
        origin, size = dec.lid_region(dims)
        o.set_flags(origin, size, orc.FLAG_INJECTION)
        s.set(_capi.FIELD_FLAGS, np.full(size[::-1], orc.FLAG_INJECTION, np.int32), origin, size)
    dd0 = o.dd.copy() if amp else None
    r_dd, r_vel, r_dens, r_fl = run_reference(dims, wg, bc, inv_tau, g, lid, o.flags, dd0, steps)
    np.testing.assert_array_equal(r_fl, o.flags)
    o.run(steps)
    s.run(steps)
    g_dd = s.store(_capi.FIELD_DD).reshape(19, -1)
    fluid = o.flags == orc.FLAG_FLUID
    r_rho, r_j = moments(r_dd, steps)
    g_rho, g_j = moments(g_dd, steps)
    o_rho, o_j = moments(o.dd, steps)
    e = dict(rho_gpu=rel(g_rho[fluid], r_rho[fluid]), j_gpu=rel(g_j[:, fluid], r_j[:, fluid]),
             rho_oracle=rel(o_rho[fluid], r_rho[fluid]), j_oracle=rel(o_j[:, fluid], r_j[:, fluid]),
             vel_gpu=rel(s.store(_capi.FIELD_VELOCITY).reshape(3, -1)[:, fluid], r_vel[:, fluid]),
             bitwise_fluid_slots=float(np.mean(g_dd[:, fluid] == r_dd[:, fluid])))
    say("  " + ", ".join(f"{k}={v:.3e}" for k, v in e.items()))
    s.close()
    for k in ("rho_gpu", "j_gpu", "rho_oracle", "j_oracle", "vel_gpu"):
        assert e[k] <= RTOL, f"{k} = {e[k]:.3e} exceeds {RTOL}"
    return e


def test_reference_config_64_cavity_1000_steps():
    """The plain reference method: 64^3 BGK cavity, tau 0.6, lid 0.05, f0 = w,
    1000 steps (the work-group size is a multiple of X: the reference's
    shortcut LDS path)."""
    compare("64^3 cavity", (64, 64, 64), 128, (1,) * 6, tau=0.6, lid=0.05, steps=1000)


@pytest.mark.parametrize("steps", [1, 2, 51])
def test_reference_general_lds_path_perturbed(steps):
    compare("20x18x16 perturbed", (20, 18, 16), 64, (1,) * 6, tau=0.6, lid=0.05, amp=1e-3, seed=20260101,
            steps=steps)


def test_reference_ghost_faces():
    """GHOST faces pass populations through (lbm_beta.cl:654-655) with the
    linear wrap of the general LDS path; without an exchange the wrapped
    values enter the fluid, identically on both sides."""
    compare("20x18x16 ghost faces", (20, 18, 16), 64, (8, 1, 1, 1, 8, 1), tau=0.7, lid=0.04, amp=1e-3, seed=5,
            steps=30)


def test_reference_forcing_alias():
    """Gravitation with g_x != 0, which the reference's alpha kernel applies
    through its aliased term chain (lbm_alpha.cl:173-298)."""
    compare("32x24x16 forcing", (32, 24, 16), 128, (1,) * 6, tau=0.7, lid=0.04, g=(2e-5, -1e-5, 3e-6), amp=1e-3,
            seed=5, steps=30)
