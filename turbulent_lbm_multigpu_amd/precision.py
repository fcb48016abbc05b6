"""Precision study: reduced-precision storage of the distributions against FP64.

Workload (the channel configuration): a D3Q19 Smagorinsky LES channel
periodic in x, walls in y and z, driven by a body force g_x, with seeded
spherical obstacles; the distributions are stored as FP32, FP16(f - w_i) or
BF16(f - w_i) with FP32 collision, and compared with the same run stored and
computed in FP64 on the GPU.  Reported per storage: MLUPS, the largest
|rho - rho_64| and |u - u_64| / U over fluid cells (U = largest |u_64|), the
relative kinetic-energy and mass errors from the FP64 reductions.

    python -m turbulent_lbm_multigpu_amd.precision [--dims 384 192 192] [--steps 10000]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import _capi
from . import decomposition as dec
from .solver import LbmGroup, LbmSolver

STORAGES = {"f64": _capi.STORAGE_F64, "f32": _capi.STORAGE_F32, "f16": _capi.STORAGE_F16_SHIFTED,
            "bf16": _capi.STORAGE_BF16_SHIFTED}
# fastest measured kernel variant per storage on this channel: non-temporal
# accesses, one cell per work-item for f64/f32, two for the two-byte storages
# (their pair of cells moves as one 32-bit word), tiles dealt to the XCDs in
# chunks (| 32)
DEFAULT_VARIANT = {"f64": 35, "f32": 35, "f16": 43, "bf16": 43}


def build_channel(dims, storage, *, tau=0.51, cs=0.17, gx=1e-6, sphere_seed=20260101, fill=0.05, rmin=3.0,
                  rmax=8.0, amp=0.0, seed=0, kernel_variant=0, work_group_size=256, fused_x=_capi.FUSED_X_AUTO,
                  alpha_work_group=0):
    """One periodic channel in a group of one subdomain whose two x comms point
    at itself.  The periodic seam is a ghost-layer exchange; with fused_x 1 or
    _capi.FUSED_X_AUTO it runs inside the step kernel: as run tiles for one
    cell per lane, as seam wavefronts for packed pairs (variant 8 or 11 of a
    two-byte storage, AUTO only)."""
    sub = dec.decompose(dims, (1, 1, 1), periodic=(True, False, False))[0]
    bc = [sub.bc[0][0], sub.bc[0][1], sub.bc[1][0], sub.bc[1][1], sub.bc[2][0], sub.bc[2][1]]
    s = LbmSolver(dims, bc, inv_tau=1.0 / tau, tau=tau, smagorinsky_cs=cs, gravitation=(gx, 0.0, 0.0),
                  forcing=_capi.FORCING_BODY, storage=storage, kernel_variant=kernel_variant,
                  work_group_size=work_group_size)
    if amp:
        s.init_perturbed(amp, seed)
    achieved = s.place_spheres(sphere_seed, fill, rmin, rmax) if fill > 0 else 0.0
    g = LbmGroup([s])
    for c in sub.comms:
        g.add_comm(0, 0, c.send_origin, c.recv_origin, c.send_size, c.comm_direction)
    g.set_exchange_mode(0)
    g.sync(after_beta=False)   # ghost layers start as copies of their periodic images
    g.set_exchange_mode(1)
    g.set_fused_x(fused_x)
    if alpha_work_group:
        g.set_alpha_work_group(alpha_work_group)
    return s, g, achieved


def fields(s: LbmSolver):
    rho, j = s.moments()
    flags = s.store(_capi.FIELD_FLAGS)
    fluid = flags == _capi.FLAG_FLUID
    rho = rho.astype(np.float64)
    u = j.astype(np.float64) / rho
    return rho, u, fluid


def run(dims=(384, 192, 192), steps=10000, storages=("f64", "f32", "f16", "bf16"), kernel_variant=None, **kw):
    """kernel_variant None: DEFAULT_VARIANT of each storage."""
    out = {}
    ref = None
    for name in storages:
        variant = DEFAULT_VARIANT[name] if kernel_variant is None else kernel_variant
        s, g, filled = build_channel(dims, STORAGES[name], kernel_variant=variant, **kw)
        m0, k0 = s.mass_energy()
        g.step(2)
        g.wait()
        t0 = time.perf_counter()
        g.step(steps - 2)
        g.wait()
        el = time.perf_counter() - t0
        m1, k1 = s.mass_energy()
        rho, u, fluid = fields(s)
        r = {"mlups": round(np.prod(dims) * (steps - 2) / el / 1e6, 1), "kernel_variant": variant,
             "mass_start": m0, "mass_end": m1,
             "kinetic_energy": k1, "obstacle_fill": round(filled, 4)}
        if name == "f64":
            ref = (rho, u, fluid, m1, k1)
        elif ref is not None:   # errors only against an FP64 run earlier in `storages`
            rr, ur, fr, mr, kr = ref
            U = np.abs(ur[:, fr]).max()
            r.update(rho_max_abs_err=float(np.abs(rho[fr] - rr[fr]).max()),
                     u_max_abs_err_over_U=float(np.abs(u[:, fr] - ur[:, fr]).max() / U),
                     u_rms_err_over_U=float(np.sqrt(np.mean((u[:, fr] - ur[:, fr]) ** 2)) / U),
                     ke_rel_err=float(abs(k1 - kr) / kr), mass_rel_err=float(abs(m1 - mr) / mr), U_ref=float(U))
        out[name] = r
        print(json.dumps({name: r}), file=sys.stderr, flush=True)
        g.close()
        s.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--dims", type=int, nargs=3, default=[384, 192, 192])
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--storages", nargs="+", default=["f64", "f32", "f16", "bf16"])
    ap.add_argument("--variant", type=int, default=None,
                    help="kernel variant for every storage (default: 35 for f64/f32, 43 for f16/bf16)")
    ap.add_argument("--fused-x", type=int, default=_capi.FUSED_X_AUTO, choices=[0, 1, _capi.FUSED_X_AUTO],
                    help="periodic x faces by a copy launch (0), inside the step kernel (1), or per "
                         "launch shape (2: run tiles for one cell per lane, seam wavefronts for packed "
                         "pairs of a two-byte storage)")
    ap.add_argument("--wg", type=int, default=256, help="work-group size (64..512, a multiple of 64)")
    ap.add_argument("--wg-alpha", type=int, default=0, help="alpha's work-group size (0: --wg)")
    a = ap.parse_args()
    res = run(tuple(a.dims), a.steps, tuple(a.storages), kernel_variant=a.variant, fused_x=a.fused_x,
              work_group_size=a.wg, alpha_work_group=a.wg_alpha)
    print(json.dumps({"workload": f"LES channel {a.dims[0]}x{a.dims[1]}x{a.dims[2]}, periodic x, g_x 1e-6, "
                                  f"tau 0.51, Cs 0.17, 5% seeded spheres r 3-8, {a.steps} steps",
                      "results": res}))


if __name__ == "__main__":
    main()
