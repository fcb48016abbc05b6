"""Writes the golden vectors of tests/test_golden_reference.py.

Runs the reference's own kernels (lbm_init.cl, lbm_alpha.cl, lbm_beta.cl,
unmodified, built into oracle/_ref by oracle/build_ref.py) through
oracle/_ref/libref_runner.so on a GPU, from initial states the C oracle
writes, and saves rho and j (from the distributions, FP64 sums stored as
float32) and the flags of each case as a small .npz.  Run on a machine with
an MI355X after building oracle/_ref:

    python tests/golden/make_reference_vectors.py --out tests/golden
"""
import argparse
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from turbulent_lbm_multigpu_amd import decomposition as dec  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")

# name: grid, work-group size (selects the LDS path of lbm_beta.cl), boundary
# flags, tau, lid, gravitation, perturbation amplitude and seed, steps
CASES = {
    "general_lds_perturbed": dict(dims=(20, 18, 16), wg=64, bc=(1,) * 6, tau=0.6, lid=0.05, g=(0.0, 0.0, 0.0),
                                  amp=1e-3, seed=20260101, steps=51),
    "ghost_faces": dict(dims=(20, 18, 16), wg=64, bc=(8, 1, 1, 1, 8, 1), tau=0.7, lid=0.04, g=(0.0, 0.0, 0.0),
                        amp=1e-3, seed=5, steps=30),
    "shortcut_lds_forcing_alias": dict(dims=(32, 24, 16), wg=128, bc=(1,) * 6, tau=0.7, lid=0.04,
                                       g=(2e-5, -1e-5, 3e-6), amp=1e-3, seed=5, steps=30),
}


def oracle_initial(case):
    """The initial domain: the oracle's perturbed state with the lid layer."""
    dims = case["dims"]
    inv_tau = float(np.float32(1.0) / np.float32(case["tau"]))
    o = orc.OracleDomain(dims, case["bc"], inv_tau=inv_tau, lid_velocity=case["lid"], gravitation=case["g"])
    o.reset(amplitude=case["amp"], seed=case["seed"])
    origin, size = dec.lid_region(dims)
    o.set_flags(origin, size, orc.FLAG_INJECTION)
    return o, inv_tau


def moments(dd, steps):
    E = orc.E.astype(np.float64)
    f = (dd[orc.OPP] if steps % 2 == 0 else dd).astype(np.float64)
    return f.sum(0), E.T @ f


def run_reference(case, o, inv_tau):
    lib = C.CDLL(os.path.join(REF_DIR, "libref_runner.so"))
    P, f = C.c_void_p, C.c_float
    lib.ref_run.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, P, f, f, f, f, f, P, P, C.c_int,
                            P, P, P, P]
    lib.ref_run.restype = C.c_int
    lib.ref_last_error.restype = C.c_char_p
    dims, wg = case["dims"], case["wg"]
    n = int(np.prod(dims))
    dd = np.zeros((19, n), np.float32)
    vel = np.zeros((3, n), np.float32)
    dens = np.zeros(n, np.float32)
    fl = np.zeros(n, np.int32)
    co = os.path.join(REF_DIR, f"ref_{dims[0]}x{dims[1]}x{dims[2]}_wg{wg}.co").encode()
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    bc = np.asarray(case["bc"], np.int32)
    flags = np.ascontiguousarray(o.flags, np.int32)
    dd0 = np.ascontiguousarray(o.dd, np.float32)
    g = case["g"]
    rc = lib.ref_run(co, dims[0], dims[1], dims[2], wg, ptr(bc), np.float32(inv_tau), np.float32(g[0]),
                     np.float32(g[1]), np.float32(g[2]), np.float32(case["lid"]), ptr(flags), ptr(dd0),
                     case["steps"], ptr(dd), ptr(vel), ptr(dens), ptr(fl))
    if rc != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    return dd, fl


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    for name, case in CASES.items():
        o, inv_tau = oracle_initial(case)
        dd, fl = run_reference(case, o, inv_tau)
        if not np.array_equal(fl, o.flags):
            raise RuntimeError(f"{name}: reference init flags differ from the oracle's")
        rho, j = moments(dd, case["steps"])
        path = os.path.join(a.out, f"ref_{name}.npz")
        np.savez_compressed(path, rho=rho.astype(np.float32), j=j.astype(np.float32), flags=fl.astype(np.uint8),
                            params=np.frombuffer(json.dumps(case).encode(), np.uint8))
        print(f"{name}: {path}", flush=True)


if __name__ == "__main__":
    main()
