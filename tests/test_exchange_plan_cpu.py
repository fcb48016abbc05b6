"""The exchange planner on the CPU (g++, no GPU): the chains of x-joined
members the step kernel exchanges in LDS, and the reference's sequential
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


def run_selftest(exe):
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    return out.returncode, out.stdout.strip().splitlines()


def test_merged_plan_equals_sequential_phases(selftest):
    rc, lines = run_selftest(selftest)
    lines = [ln for ln in lines if not ln.startswith("chains ")]
    assert rc == 0, "\n".join(ln for ln in lines if "MISMATCH" in ln)
    assert len(lines) == 7 * 4 * 2 * 2 * 2
    assert all(ln.endswith("merged ok") for ln in lines), [ln for ln in lines if not ln.endswith("merged ok")]


def test_chain_planning(selftest):
    """Joins (conflicting and one-sided ones dropped), lines and rings in x
    order, chains longer than 8 left out, and the chain-tile shapes (whole
    rows, or face runs plus interior tiles for wide rows)."""
    rc, lines = run_selftest(selftest)
    chains = [ln for ln in lines if ln.startswith("chains ")]
    assert len(chains) == 14
    assert all(ln.endswith(" ok") for ln in chains), [ln for ln in chains if not ln.endswith(" ok")]
