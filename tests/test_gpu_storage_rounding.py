"""The two-byte storages round f - w_i to nearest even, bit for bit.

Writing distributions through LBM_FIELD_DD converts them on the device to
the stored FP16 / BF16 bits of f - w_i (the weights of the FP32 path,
src/cl_programs/lbm_header.h:68-94). The stored bits are read back through
LBM_FIELD_DD_RAW and compared with numpy's restatement: IEEE binary16
conversion (numpy's float16 cast rounds to nearest even), and the integer
round-to-nearest-even u + 0x7fff + ((u >> 16) & 1) for bfloat16. The inputs
cover ties (exactly half an ulp away), values that round across a binade,
FP16 subnormals and both signs. Tolerance 0.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from turbulent_lbm_multigpu_amd.solver import LbmSolver  # noqa: E402

F = np.float32
W = np.array([F(1) / F(18) if (i < 4 or i >= 16) else F(1) / F(36) for i in range(18)] + [F(1) / F(3)], F)
DIMS = (32, 16, 16)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bf16_rne(d):
    u = d.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def shifted_inputs(storage, seed):
    """Distributions f (float32) whose f - w_i hit the rounding cases of the
    storage; returns f and the exact float32 f - w_i."""
    rng = np.random.default_rng(seed)
    n = DIMS[0] * DIMS[1] * DIMS[2]
    d = np.empty((19, n), F)
    quarter = n // 4
    # random magnitudes over many binades, both signs
    mag = np.exp2(rng.uniform(-26 if storage == _capi.STORAGE_F16_SHIFTED else -40, 1, (19, quarter)))
    d[:, :quarter] = (mag * rng.choice([-1.0, 1.0], (19, quarter))).astype(F)
    # exact ties: the bits below the kept mantissa are exactly one half
    keep = 16 if storage == _capi.STORAGE_BF16_SHIFTED else 13
    base = d[:, :quarter].view(np.uint32) & np.uint32(~((1 << keep) - 1) & 0xFFFFFFFF)
    d[:, quarter:2 * quarter] = (base | np.uint32(1 << (keep - 1))).view(F)
    # just below and just above a tie
    d[:, 2 * quarter:3 * quarter] = (base | np.uint32((1 << (keep - 1)) - 1)).view(F)
    d[:, 3 * quarter:] = (base[:, : n - 3 * quarter] | np.uint32((1 << (keep - 1)) + 1)).view(F)
    if storage == _capi.STORAGE_F16_SHIFTED:
        # keep |d| below FP16's largest finite value
        d = np.clip(d, F(-4), F(4))
    f = (d + W[:, None]).astype(F)
    return f, (f - W[:, None]).astype(F)


@pytest.mark.parametrize("storage", [_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
@pytest.mark.parametrize("seed", [3, 4])
def test_stored_bits_round_to_nearest_even(storage, seed):
    say(f"case: storage rounding {storage} seed {seed}")
    f, d = shifted_inputs(storage, seed)
    if storage == _capi.STORAGE_BF16_SHIFTED:
        want = bf16_rne(d)
        ties = np.count_nonzero((d.view(np.uint32) & 0xFFFF) == 0x8000)
        back = (want.astype(np.uint32) << 16).view(F)
    else:
        want = d.astype(np.float16).view(np.uint16)
        back = want.view(np.float16).astype(F)
        ties = np.count_nonzero((d.view(np.uint32) & 0x1FFF) == 0x1000)
        subnormal = np.count_nonzero((want & 0x7C00) == 0)
        say(f"  results in FP16 subnormal range: {subnormal}")
        assert subnormal > 1000
    say(f"  exact ties among f - w_i: {ties}")
    assert ties > 1000
    s = LbmSolver(DIMS, inv_tau=1 / 0.6, tau=0.6, storage=storage)
    try:
        s.set(_capi.FIELD_DD, f.reshape(19, DIMS[2], DIMS[1], DIMS[0]))
        raw = s.store(_capi.FIELD_DD_RAW).reshape(19, -1)
        assert raw.dtype == np.uint16
        np.testing.assert_array_equal(raw, want)
        # reading back: the stored value plus w_i in FP32
        np.testing.assert_array_equal(s.store(_capi.FIELD_DD).reshape(19, -1), (back + W[:, None]).astype(F))
    finally:
        s.close()
