"""Checkpoint and resume through the full-state get/set (distributions,
flags, step counter): a run interrupted, saved, restored into a fresh solver
and continued is bit-identical to the uninterrupted run (tolerance 0)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from turbulent_lbm_multigpu_amd import decomposition as dec  # noqa: E402
from turbulent_lbm_multigpu_amd.solver import LbmSolver  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cavity(storage):
    dims = (22, 18, 20)
    s = LbmSolver(dims, (1,) * 6, inv_tau=1 / 0.53, tau=0.53, smagorinsky_cs=0.17, lid_velocity=0.07, storage=storage)
    s.init_perturbed(1e-3, 31)
    o, sz = dec.lid_region(dims)
    s.set(_capi.FIELD_FLAGS, np.full(sz[::-1], _capi.FLAG_VELOCITY_INJECTION, np.int32), o, sz)
    return s


@pytest.mark.parametrize("storage", [_capi.STORAGE_F32, _capi.STORAGE_F64, _capi.STORAGE_F16_SHIFTED,
                                     _capi.STORAGE_BF16_SHIFTED])
@pytest.mark.parametrize("first", [10, 7])
def test_checkpoint_resume_bitwise(storage, first):
    say(f"case: checkpoint storage {storage} after {first} steps")
    total = 23
    ref = cavity(storage)
    ref.run(total)
    want = ref.store(_capi.FIELD_DD)
    ref.close()
    a = cavity(storage)
    a.run(first)
    state = a.checkpoint()
    a.close()
    b = cavity(storage)
    b.reset()                      # a fresh state, then the checkpoint over it
    b.restore(state)
    assert b.step_counter == first
    b.run(total - first)
    np.testing.assert_array_equal(b.store(_capi.FIELD_DD), want)
    b.close()


@pytest.mark.parametrize("storage", [_capi.STORAGE_F32, _capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
def test_raw_field_roundtrip_and_region(storage):
    """LBM_FIELD_DD_RAW moves the stored bits: a region set from a store is a
    no-op, and the raw bits of FP32 storage are the FP32 field's bits."""
    say(f"case: raw field storage {storage}")
    s = cavity(storage)
    s.run(5)
    raw = s.store(_capi.FIELD_DD_RAW)
    o, sz = (3, 2, 4), (7, 5, 6)
    part = s.store(_capi.FIELD_DD_RAW, o, sz)
    np.testing.assert_array_equal(part, raw[:, 4:10, 2:7, 3:10])
    s.set(_capi.FIELD_DD_RAW, part, o, sz)
    np.testing.assert_array_equal(s.store(_capi.FIELD_DD_RAW), raw)
    if storage == _capi.STORAGE_F32:
        np.testing.assert_array_equal(raw.view(np.float32), s.store(_capi.FIELD_DD))
    s.close()
