"""Build recipe of the CPU oracle (test infrastructure only).

Compiles oracle/lbm_oracle.c with gcc -O3 -fopenmp -ffp-contract=off into
oracle/liblbm_oracle.so.  -ffp-contract=off keeps every multiply and add a
separate IEEE operation, as in the HIP library.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lbm_oracle.c")
LIB = os.path.join(HERE, "liblbm_oracle.so")


def build(verbose: bool = False) -> str:
    gcc = shutil.which("gcc") or "gcc"
    cmd = [gcc, "-O3", "-march=x86-64-v2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-Wall", "-o", LIB + ".tmp", SRC, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
