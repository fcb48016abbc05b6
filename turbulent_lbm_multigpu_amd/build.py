"""Build recipe of the HIP library (gfx950) and the C++ host tools.

``build_library()`` compiles ``csrc/*.hip`` and ``csrc/*.cpp`` into
``liblbm_hip.so`` next to this file (in-tree, so it travels with the source
tree).  ``build_tools()`` compiles the C++ command-line driver that uses the
header-only host layer in ``csrc/host``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "liblbm_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build the library)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        ln for ln in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in ln
    )
    if res.returncode != 0:
        raise RuntimeError(f"build failed ({' '.join(cmd[:2])} ...):\n{out}")
    if verbose and out.strip():
        print(out)


def library_sources():
    srcs = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith(".hip") or (name.endswith(".cpp") and name != "lbm_main.cpp"):
            srcs.append(os.path.join(CSRC, name))
    return srcs


def build_library(verbose: bool = False) -> str:
    """Compile the HIP runtime into LIB_PATH; returns the path."""
    hipcc = _hipcc()
    objs = []
    tmpdir = os.path.join(HERE, "_build")
    os.makedirs(tmpdir, exist_ok=True)
    for src in library_sources():
        obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
        cmd = [hipcc, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-x", "hip", *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj]
        _run(cmd, verbose)
        objs.append(obj)
    tmp_lib = LIB_PATH + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_lib, *objs], verbose)
    os.replace(tmp_lib, LIB_PATH)
    return LIB_PATH


def build_tools(verbose: bool = False) -> str:
    """Compile the C++ driver (csrc/lbm_main.cpp) against the library."""
    main_src = os.path.join(CSRC, "lbm_main.cpp")
    if not os.path.exists(main_src):
        return ""
    out = os.path.join(HERE, "lbm_main")
    cmd = [
        shutil.which("g++") or "g++", "-O2", "-std=c++17", "-Wall",
        f"-I{INCLUDE}", f"-I{CSRC}", main_src, "-o", out,
        f"-L{HERE}", "-llbm_hip", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN",
    ]
    _run(cmd, verbose)
    return out


if __name__ == "__main__":
    v = "-v" in sys.argv
    print(build_library(verbose=v))
    t = build_tools(verbose=v)
    if t:
        print(t)
