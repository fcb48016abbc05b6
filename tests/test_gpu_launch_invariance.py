"""Results do not depend on the launch shape.

Seeded random groups (subdomain sizes, decompositions, periodic axes,
storage, exchange mode, a frozen member) are stepped twice from the same
state: once with one plain shape (work-group 256, kernel variant 0, x faces by
the copy launch, graph replay) and once with a random shape (work-group size,
kernel variant with its cache policy, cells per lane and XCD tile map, alpha's own
work-group size, x faces inside the kernel or by copies, graph or eager
launches, subdomains in one allocation or not; a fifth of the cases are
periodic rings of one in packed pairs, which take the seam wavefronts).
Every slot of every
subdomain, the velocity and the density outputs must be bit-identical after
odd and after even step counts: every launch shape runs the same per-cell
arithmetic, so any difference is an indexing, exchange or tiling error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from tests.test_gpu_parity import build_gpu_group, gpu_dd, say  # noqa: E402

NUMS = [(1, 1, 1), (2, 1, 1), (3, 1, 1), (2, 2, 1), (1, 2, 2), (2, 1, 2), (2, 2, 2), (4, 1, 1)]
STORAGES = [_capi.STORAGE_F32, _capi.STORAGE_F64, _capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED]
VARIANTS = [0, 1, 2, 3, 8, 11, 16]
XCD_MAPS = [0, 0, 4, 32]   # eighths (twice as likely), none, chunks
WGS = [64, 128, 256, 512]


def draw(case: int):
    """One random group and launch shape, all within the documented limits."""
    rng = np.random.default_rng(20261203 + case)
    nums = NUMS[int(rng.integers(len(NUMS)))]
    sub = []
    for a in range(3):
        hi = 40 if nums[a] > 1 else 48
        if a == 0 and nums[a] == 1 and rng.random() < 0.25:
            hi = 200   # long rows: run tiles, and packed-pair seams when periodic
        sub.append(int(rng.integers(6, hi)))
    dims = tuple(sub[a] * nums[a] for a in range(3))
    periodic = tuple(bool(rng.random() < 0.3) for _ in range(3))
    shape = dict(
        wg=WGS[int(rng.integers(4))],
        variant=VARIANTS[int(rng.integers(len(VARIANTS)))] | XCD_MAPS[int(rng.integers(len(XCD_MAPS)))],
        alpha_wg=[0, 64, 128, 256, 512][int(rng.integers(5))],
        fused=[0, 1, _capi.FUSED_X_AUTO][int(rng.integers(3))],
        graph=bool(rng.random() < 0.7),
        colocated=bool(rng.random() < 0.3),
    )
    storage = STORAGES[int(rng.integers(4))]
    if rng.random() < 0.2:
        # a periodic ring of one with rows of an even length >= 128 in packed
        # pairs: the seam wavefronts
        nums = (1, int(rng.integers(1, 3)), 1)
        dims = (2 * int(rng.integers(64, 100)), nums[1] * int(rng.integers(6, 16)), int(rng.integers(6, 16)))
        periodic = (True, periodic[1], periodic[2])
        storage = STORAGES[2 + int(rng.integers(2))]
        shape.update(variant=(8, 11)[int(rng.integers(2))] | XCD_MAPS[int(rng.integers(len(XCD_MAPS)))],
                     fused=(1, _capi.FUSED_X_AUTO)[int(rng.integers(2))])
    n = nums[0] * nums[1] * nums[2]
    stop = (int(rng.integers(n)), int(rng.integers(1, 6))) if rng.random() < 0.3 else None
    return dict(dims=dims, nums=nums, periodic=periodic, storage=storage, mode=int(rng.integers(2)), shape=shape,
                stop=stop)


@pytest.mark.parametrize("case", range(64))
def test_launch_shape_does_not_change_results(case):
    c = draw(case)
    sh = c["shape"]
    say(f"case {case}: dims {c['dims']} nums {c['nums']} periodic {c['periodic']} storage {c['storage']} "
        f"mode {c['mode']} stop {c['stop']} shape {sh}")
    kw = dict(inv_tau=1 / 0.52, tau=0.52, smagorinsky_cs=0.17, storage=c["storage"])
    common = dict(lid=0.08, amp=1e-3, seed=7 + case, mode=c["mode"], periodic=c["periodic"])
    ref, _, _ = build_gpu_group(c["dims"], c["nums"], graph=True, **common, **kw)
    tst, _, _ = build_gpu_group(c["dims"], c["nums"], graph=sh["graph"], colocated=sh["colocated"], **common, **kw)
    tst.set_launch(sh["wg"], sh["variant"])
    tst.set_alpha_work_group(sh["alpha_wg"])
    tst.set_fused_x(sh["fused"])
    if c["stop"] is not None:
        for g in (ref, tst):
            g.set_member_stop(*c["stop"])
    for steps in (3, 4):
        for g in (ref, tst):
            g.step(steps)
            g.wait()
        for k, (a, b) in enumerate(zip(ref.solvers, tst.solvers)):
            np.testing.assert_array_equal(gpu_dd(b), gpu_dd(a), err_msg=f"member {k} distributions")
            np.testing.assert_array_equal(b.store(_capi.FIELD_VELOCITY), a.store(_capi.FIELD_VELOCITY),
                                          err_msg=f"member {k} velocity")
            np.testing.assert_array_equal(b.store(_capi.FIELD_DENSITY), a.store(_capi.FIELD_DENSITY),
                                          err_msg=f"member {k} density")
    for g in (ref, tst):
        g.close()
        for s in g.solvers[::-1]:
            s.close()
