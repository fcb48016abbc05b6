"""The C oracle against golden vectors from the reference's own kernels.

tests/golden/ref_*.npz hold rho and j after N steps of lbm_init.cl,
lbm_alpha.cl and lbm_beta.cl (unmodified, built for gfx950 and run by
tests/golden/make_reference_vectors.py) from initial states the oracle wrote.
The oracle must reproduce them within a relative 1e-5 on FLUID cells (the
reference's OpenCL build may contract multiply-adds, so bit equality is not
expected); flags must match exactly.  No GPU is needed.
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.golden.make_reference_vectors import moments, oracle_initial

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "ref_*.npz")))
RTOL = 1e-5


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.skipif(not FILES, reason="no golden vectors committed")
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(p) for p in FILES])
def test_oracle_matches_reference_kernels(path):
    z = np.load(path, allow_pickle=False)
    case = json.loads(bytes(z["params"]).decode())
    case["dims"], case["bc"], case["g"] = tuple(case["dims"]), tuple(case["bc"]), tuple(case["g"])
    o, _ = oracle_initial(case)
    np.testing.assert_array_equal(o.flags, z["flags"])
    o.run(case["steps"])
    rho, j = moments(o.dd, case["steps"])
    fluid = z["flags"] == orc.FLAG_FLUID
    assert fluid.sum() > 100
    assert rel(rho[fluid], z["rho"][fluid].astype(np.float64)) <= RTOL
    assert rel(j[:, fluid], z["j"][:, fluid].astype(np.float64)) <= RTOL


@pytest.mark.skipif(not FILES, reason="no golden vectors committed")
def test_golden_vectors_detect_a_planted_error():
    """The comparison has teeth: the oracle with tau or the lid speed off by a
    small relative amount, or read out in the wrong step parity (slot i vs
    opp(i)), misses the reference's fields by far more than 1e-5."""
    path = [p for p in FILES if "general_lds" in p][0]
    z = np.load(path, allow_pickle=False)
    case = json.loads(bytes(z["params"]).decode())
    case["dims"], case["bc"], case["g"] = tuple(case["dims"]), tuple(case["bc"]), tuple(case["g"])
    fluid = z["flags"] == orc.FLAG_FLUID
    for planted in ("tau", "lid", "parity"):
        c = dict(case)
        if planted == "tau":
            c["tau"] = case["tau"] * (1 + 1e-4)
        if planted == "lid":
            c["lid"] = case["lid"] * (1 + 1e-3)
        o, _ = oracle_initial(c)
        o.run(case["steps"])
        rho, j = moments(o.dd, case["steps"] + (1 if planted == "parity" else 0))
        assert rel(j[:, fluid], z["j"][:, fluid].astype(np.float64)) > 10 * RTOL, planted
