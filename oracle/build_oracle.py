"""Build recipes of the CPU oracle (test infrastructure only).

``build()`` compiles oracle/lbm_oracle.c with gcc -O3 -fopenmp
-ffp-contract=off into oracle/liblbm_oracle.so, the parity build: every
multiply and add stays a separate IEEE operation, as in the HIP library.

``build_timing()`` compiles the same source for speed (-O3 -march=native,
contraction allowed) into a directory of the caller's choice.  It is the
timed CPU baseline only, never a parity checker, and is built on the machine
that runs it, because -march=native targets that machine's own CPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lbm_oracle.c")
LIB = os.path.join(HERE, "liblbm_oracle.so")

PARITY_FLAGS = ["-O3", "-march=x86-64-v2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                "-Wall"]
TIMING_FLAGS = ["-O3", "-march=native", "-fPIC", "-shared", "-fopenmp", "-Wall"]


def _gcc(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)


def build(verbose: bool = False) -> str:
    gcc = shutil.which("gcc") or "gcc"
    _gcc([gcc, *PARITY_FLAGS, "-o", LIB + ".tmp", SRC, "-lm"], verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_timing(out_dir: str, verbose: bool = False) -> str:
    """The timing build in out_dir/liblbm_oracle_timing.so; returns its path."""
    gcc = shutil.which("gcc") or "gcc"
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "liblbm_oracle_timing.so")
    _gcc([gcc, *TIMING_FLAGS, "-o", out + ".tmp", SRC, "-lm"], verbose)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
