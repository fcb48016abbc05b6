"""Batching on the GPU: many uneven cases in one fused launch.

Every case's result must be bit-identical to the same case run on its own:
the per-cell arithmetic does not depend on the launch shape, on which tiles
share a launch, or on members joining, leaving or freezing (tolerance 0).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from turbulent_lbm_multigpu_amd import batch  # noqa: E402
from turbulent_lbm_multigpu_amd.solver import LbmGroup, LbmSolver  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dd(s):
    return s.store(_capi.FIELD_DD)


def alone(c, steps, **kw):
    s = batch.build_solver(c, **kw)
    s.run(steps)
    s.wait()
    out = dd(s), s.step_counter
    s.close()
    return out


CASES = [batch.Cavity((17, 21, 19), 0.08, 0.51), batch.Cavity((24, 16, 30), 0.05, 0.5003),
         batch.Cavity((33, 20, 18), 0.1, 0.55), batch.Cavity((16, 16, 16), 0.03, 0.6)]


@pytest.mark.parametrize("storage", [_capi.STORAGE_F32, _capi.STORAGE_F16_SHIFTED])
def test_member_stops_match_individual_runs(storage):
    """Members frozen at different step counts (odd ones included) inside one
    graph-replayed sequence equal their individual runs of that length."""
    say(f"case: member stops, storage {storage}")
    stops = [7, 12, 20, 1]
    sol = [batch.build_solver(c, storage=storage) for c in CASES]
    g = LbmGroup.empty(storage)
    for s, k in zip(sol, stops):
        i = g.add_member(s)
        g.set_member_stop(i, k)
    g.step(24)
    g.wait()
    for s, c, k in zip(sol, CASES, stops):
        assert s.step_counter == k
        ref, n = alone(c, k, storage=storage)
        assert n == k
        np.testing.assert_array_equal(dd(s), ref)
    g.close()
    for s in sol:
        s.close()


def test_join_and_leave_between_chunks():
    """A case joins an even-counted group mid-run and another leaves; both end
    bit-identical to individual runs of their own step counts."""
    say("case: join/leave")
    a, b, c = (batch.build_solver(x) for x in CASES[:3])
    g = LbmGroup.empty(_capi.STORAGE_F32, work_group_size=128, kernel_variant=8)
    g.add_member(a)
    g.add_member(b)
    g.step(10)
    g.remove_member(0)          # a leaves after 10 steps; b renumbered to 0
    g.add_member(c)             # c joins at group step 10 (even)
    g.set_member_stop(1, 9)     # ... and stops after 9 of its own steps
    g.step(14)
    g.wait()
    assert (a.step_counter, b.step_counter, c.step_counter) == (10, 24, 9)
    for s, x, k in ((a, CASES[0], 10), (b, CASES[1], 24), (c, CASES[2], 9)):
        ref, _ = alone(x, k)
        np.testing.assert_array_equal(dd(s), ref)
    g.close()
    for s in (a, b, c):
        s.close()


def test_periodic_channel_member_freeze():
    """A periodic channel (self comms across x) frozen mid-chunk keeps its
    exchanges frozen too: equal to the channel run alone for its steps."""
    say("case: channel freeze")
    from turbulent_lbm_multigpu_amd import precision

    ref_s, ref_g, _ = precision.build_channel((24, 18, 16), _capi.STORAGE_F32, tau=0.55, gx=2e-5, fill=0.04,
                                              rmin=2.0, rmax=3.0, amp=1e-3, seed=9)
    ref_g.step(13)
    ref_g.wait()
    ref = dd(ref_s)
    ref_g.close()
    ref_s.close()
    s, g, _ = precision.build_channel((24, 18, 16), _capi.STORAGE_F32, tau=0.55, gx=2e-5, fill=0.04, rmin=2.0,
                                      rmax=3.0, amp=1e-3, seed=9)
    other = batch.build_solver(CASES[0])
    g.add_member(other)
    g.set_member_stop(0, 13)
    g.step(30)
    g.wait()
    assert s.step_counter == 13 and other.step_counter == 30
    np.testing.assert_array_equal(dd(s), ref)
    ref_o, _ = alone(CASES[0], 30)
    np.testing.assert_array_equal(dd(other), ref_o)
    g.close()
    s.close()
    other.close()


def test_batched_cavities_equal_sequential():
    """The batching workload at reduced size: batched (autotuned, restarted,
    then every step in one graph launch) and one-at-a-time runs give
    bit-identical moments."""
    say("case: 12 cavities batched vs sequential, 60 steps")
    cases = batch.make_cavities(12, seed=5, edge=(16, 30))
    rb, fb = batch.run_batched(cases, 60, autotune_pairs=1, keep=True)
    rs, fs = batch.run_sequential(cases, 60, keep=True)
    say(f"  batched {rb['launch']}, sequential {rs['cases_per_s']} cases/s")
    assert rb["autotune"] is not None  # the tuner's 16 + 3 + 1 shapes ran, then the cases restarted
    assert rb["launch"]["alpha_work_group"] == rb["autotune"]["alpha_work_group"] in (64, 128, 256, 512)
    assert rb["graph_launches"] == 1 and rb["eager_half_steps"] == 0   # 60 steps in one graph launch
    for (r1, j1), (r2, j2) in zip(fb, fs):
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(j1, j2)


def test_batched_keeps_the_tuners_steps():
    """The autotuner's steps are simulation steps of every case: with more
    steps asked for than tuning takes, the batch keeps them and runs the
    rest in one graph launch, bit-identical to the cases run one at a time."""
    say("case: 10 cavities, 120 steps, tuning steps kept")
    cases = batch.make_cavities(10, seed=9, edge=(16, 28))
    rb, fb = batch.run_batched(cases, 120, autotune_pairs=1, keep=True)
    rs, fs = batch.run_sequential(cases, 120, keep=True)
    say(f"  kept {rb['tuning_steps_kept']} tuning steps, launch {rb['launch']}")
    assert 0 < rb["tuning_steps_kept"] < 120 and rb["tuning_steps_kept"] % 2 == 0
    assert rb["graph_launches"] == 1 and rb["eager_half_steps"] == 0
    for (r1, j1), (r2, j2) in zip(fb, fs):
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(j1, j2)


def test_full_size_batch_equals_sequential():
    """The batching workload at its full size (the benchmark's 512 seeded
    cavities, edges 16..96): batched with the autotuner's shapes and one at a
    time give bit-identical moments for every case after 60 steps."""
    say("case: 512 cavities batched vs sequential, 60 steps")
    cases = batch.make_cavities(512, 20260101)
    rb, fb = batch.run_batched(cases, 60, autotune_pairs=1, keep=True)
    rs, fs = batch.run_sequential(cases, 60, keep=True)
    say(f"  batched {rb['launch']}, {rb['cells']} cells")
    assert rb["autotune"] is not None
    assert rb["graph_launches"] == 1 and rb["eager_half_steps"] == 0
    for (r1, j1), (r2, j2) in zip(fb, fs):
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(j1, j2)
