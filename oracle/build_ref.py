"""Build recipe of oracle/_ref (test infrastructure only).

Compiles the reference's own OpenCL kernels -- lbm_init.cl, lbm_alpha.cl and
lbm_beta.cl, unmodified, from /root/reference/src/cl_programs -- with clang's
OpenCL C front end into gfx950 code objects, one per grid configuration.  The
-D values are the program prefix the reference host generates at run time
(src/CLbmSolver.hpp:319-373,419-422): T and its vector types, DOMAIN_CELLS_*,
GLOBAL_WORK_GROUP_SIZE, the flag values, STORE_VELOCITY/STORE_DENSITY and
LOCAL_WORK_GROUP_SIZE.  It also compiles oracle/ref_runner.hip, a launcher that
loads such a code object and replays the reference host's launch sequence.

Outputs go to oracle/_ref/ only (git-ignored, shipped with the tree to the GPU
machine).  Nothing here runs without /root/reference at build time.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
REF = "/root/reference"
CLANG = "/opt/rocm/lib/llvm/bin/clang"
LLD = "/opt/rocm/lib/llvm/bin/ld.lld"

# (nx, ny, nz, work-group size): cells must be a multiple of the work-group
# size, because the reference's init kernel has no bound check.
CONFIGS = [
    (20, 18, 16, 64),     # general LDS path of lbm_beta.cl (64 % 20 != 0)
    (64, 64, 64, 128),    # the 64^3 cavity; WG multiple of X: the shortcut path
    (32, 24, 16, 128),    # shortcut path with X = 32
]


def co_name(nx, ny, nz, wg):
    return f"ref_{nx}x{ny}x{nz}_wg{wg}.co"


def defines(nx, ny, nz, wg):
    return [
        "-DT=float", "-DT2=float2", "-DT4=float4", "-DT8=float8", "-DT16=float16",
        "-DSIZE_DD_HOST_BYTES=76",
        f"-DDOMAIN_CELLS_X={nx}", f"-DDOMAIN_CELLS_Y={ny}", f"-DDOMAIN_CELLS_Z={nz}",
        "-DGLOBAL_WORK_GROUP_SIZE=(DOMAIN_CELLS_X*DOMAIN_CELLS_Y*DOMAIN_CELLS_Z)",
        "-DFLAG_OBSTACLE=1", "-DFLAG_FLUID=2", "-DFLAG_VELOCITY_INJECTION=4", "-DFLAG_GHOST_LAYER=8",
        "-DSTORE_VELOCITY=1", "-DSTORE_DENSITY=1",
        f"-DLOCAL_WORK_GROUP_SIZE={wg}",
    ]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference build step failed: {' '.join(cmd[:3])} ...\n{r.stdout}{r.stderr}")


def build(verbose: bool = False) -> str:
    if not os.path.isdir(os.path.join(REF, "src", "cl_programs")):
        raise FileNotFoundError("reference sources not present: oracle/_ref cannot be built here")
    os.makedirs(OUT, exist_ok=True)
    tmp = os.path.join(OUT, "obj")
    os.makedirs(tmp, exist_ok=True)
    for nx, ny, nz, wg in CONFIGS:
        if (nx * ny * nz) % wg:
            raise ValueError("cells must be a multiple of the work-group size")
        objs = []
        for k in ("lbm_init", "lbm_alpha", "lbm_beta"):
            obj = os.path.join(tmp, f"{k}_{nx}x{ny}x{nz}_{wg}.o")
            _run([CLANG, "-x", "cl", "-cl-std=CL1.2", "--target=amdgcn-amd-amdhsa", "-mcpu=gfx950", "-O3",
                  "-Xclang", "-finclude-default-header", f"-I{REF}", *defines(nx, ny, nz, wg),
                  "-c", os.path.join(REF, "src", "cl_programs", f"{k}.cl"), "-o", obj], verbose)
            objs.append(obj)
        _run([LLD, "-shared", *objs, "-o", os.path.join(OUT, co_name(nx, ny, nz, wg))], verbose)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    _run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-fPIC", "-shared",
          os.path.join(HERE, "ref_runner.hip"), "-o", os.path.join(OUT, "libref_runner.so")], verbose)
    shutil.rmtree(tmp, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
