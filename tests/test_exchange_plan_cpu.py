"""The exchange planner on the CPU (g++, no GPU): the reference's sequential
x, y, z ghost-layer copy phases (src/CController.hpp:265-383) against the
merged plan that runs them as one launch, on every slot of block
decompositions built like CManager::initSimulation (src/CManager.hpp:78-204),
periodic or not, exchange modes 0 and 1, after alpha and after beta, with
and without the x faces inside the step kernel.

Every case must merge (one launch per half-step) and give bit-identical
memory; the planner's own conflict check would otherwise fall back to one
launch per axis, which the benchmark relies on not happening.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "turbulent_lbm_multigpu_amd", "csrc")


@pytest.fixture(scope="module")
def selftest(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("plan") / "exchange_plan_selftest")
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", f"-I{CSRC}",
                    os.path.join(ROOT, "tests", "cpp", "exchange_plan_selftest.cpp"), "-o", exe], check=True)
    return exe


def test_merged_plan_equals_sequential_phases(selftest):
    out = subprocess.run([selftest], capture_output=True, text=True, timeout=120)
    lines = out.stdout.strip().splitlines()
    assert out.returncode == 0, "\n".join(ln for ln in lines if "MISMATCH" in ln)
    assert len(lines) == 7 * 4 * 2 * 2 * 2
    assert all(ln.endswith("merged ok") for ln in lines), [ln for ln in lines if not ln.endswith("merged ok")]
