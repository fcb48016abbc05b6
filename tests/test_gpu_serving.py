"""The continuous-batching server on the GPU.

Stated tolerance: 0.  A case's rho and j are bit-identical whether it ran
alone or in a mix of other cases, storages and sizes admitted and retired
around it, and equal to the plain solver path of the same case.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from turbulent_lbm_multigpu_amd import batch  # noqa: E402
from turbulent_lbm_multigpu_amd import serving  # noqa: E402
from turbulent_lbm_multigpu_amd._capi import LbmError  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mix_requests():
    """Small cases arriving over ~50 ms: cavities and channels, FP32 and
    FP16, odd and even step counts, some with a perturbed start."""
    R = serving.Request
    return [
        R(0.000, "cavity", 20, 37, "f32", 11, lid=0.08),
        R(0.000, "channel", 24, 50, "f16", 12),
        R(0.005, "cavity", 28, 64, "f16", 13, lid=0.04),
        R(0.010, "channel", 18, 23, "f32", 14),
        R(0.020, "cavity", 16, 90, "f32", 15, lid=0.1),
        R(0.030, "cavity", 22, 41, "bf16", 16),
        R(0.050, "channel", 26, 60, "f32", 17),
    ]


def run_mix(reqs, **kw):
    with serving.Server(2e9, chunk_ms=0.5, max_chunk_steps=16, **kw) as srv:
        res = serving.serve(srv, reqs, keep_results=True)
    return res


def test_alone_equals_in_the_mix():
    say("case: serving mix of 7 cases vs each alone")
    reqs = mix_requests()
    mix = run_mix(reqs)
    assert mix.completed == len(reqs)
    assert mix.stats["chunks"] > 3
    for k, r in enumerate(reqs):
        one = run_mix([serving.Request(0.0, r.kind, r.edge, r.steps, r.storage, r.seed, lid=r.lid)], max_active=1)
        rho_m, j_m = mix.results[k]
        rho_a, j_a = one.results[0]
        np.testing.assert_array_equal(rho_m, rho_a)
        np.testing.assert_array_equal(j_m, j_a)
        say(f"  case {k} ({r.kind} {r.edge}^3 {r.storage} {r.steps} steps) identical")


def test_served_cavity_equals_plain_solver():
    """Server result of a cavity == the plain solver path run for the same steps."""
    say("case: served cavity vs plain solver")
    r = serving.Request(0.0, "cavity", 21, 33, "f32", 3, lid=0.07, tau=0.53)
    res = run_mix([r])
    s = batch.build_solver(batch.Cavity((21, 21, 21), 0.07, 0.53))
    s.run(33)
    rho, j = s.moments()
    s.close()
    np.testing.assert_array_equal(res.results[0][0], rho)
    np.testing.assert_array_equal(res.results[0][1], j)


def test_budget_admission_and_rejection():
    say("case: memory budget admission")
    big = serving.make_case("cavity", (40, 40, 40), 20, "f32")
    per = serving.case_bytes(big)
    budget = int(2.5 * per)  # two resident at a time
    reqs = [serving.Request(0.0, "cavity", 40, 20 + 2 * k, "f32", 100 + k) for k in range(5)]
    with serving.Server(budget, chunk_ms=0.5, max_chunk_steps=8) as srv:
        with pytest.raises(LbmError):
            srv.submit(serving.make_case("cavity", (64, 64, 64), 10, "f32"))  # can never fit
        res = serving.serve(srv, reqs)
        st = srv.stats()
    assert res.completed == 5
    assert st["bytes_peak"] <= budget
    assert st["bytes_in_use"] == 0
    assert st["steps_cells"] == sum(40 ** 3 * r.steps for r in reqs)
    say(f"  {res.summary()}")


def test_zero_step_case_returns_initial_state():
    with serving.Server(1e9, chunk_ms=0.5) as srv:
        cid = srv.submit(serving.make_case("cavity", (12, 12, 12), 0, "f32"))
        for _ in range(4):
            srv.progress()
        st = srv.status(cid)
        assert st["state"] == _capi.CASE_DONE
        rho, j = srv.fetch(cid)
    np.testing.assert_allclose(rho, 1.0, rtol=1e-6)
    assert np.abs(j).max() < 1e-6


def test_failed_case_does_not_stop_the_others():
    """A case whose setup fails (an infinite perturbation amplitude) is
    reported FAILED with the reason, its memory returns to the arena, and the
    cases around it complete."""
    say("case: failed case in the mix")
    with serving.Server(1e9, chunk_ms=0.5, max_chunk_steps=8) as srv:
        good = [srv.submit(serving.make_case("cavity", (14, 14, 14), 12 + k, "f32", seed=k)) for k in range(2)]
        bad = srv.submit(serving.make_case("cavity", (14, 14, 14), 10, "f32", perturbation=float("inf")))
        done = {}
        for _ in range(200):
            srv.progress()
            while True:
                st = srv.next_done()
                if st is None:
                    break
                done[st["id"]] = st["state"]
            if len(done) == 3:
                break
        assert done[bad] == _capi.CASE_FAILED
        assert all(done[g] == _capi.CASE_DONE for g in good)
        with pytest.raises(LbmError, match="finite"):
            srv.fetch(bad)
        for g in good:
            rho, _ = srv.fetch(g)
            assert np.isfinite(rho).all()
        st = srv.stats()
        assert st["failed"] == 1 and st["completed"] == 2 and st["bytes_in_use"] == 0


def test_full_size_mix_alone_equals_in_the_mix():
    """The serving workload at its full size (the benchmark's 24 seeded
    requests: cavities and channels of 32^3..192^3, 500..5000 steps), FP16
    storage: every case's rho and j are bit-identical served in the mix and
    served alone."""
    say("case: 24 full-size requests, FP16, in the mix vs alone")
    reqs = serving.request_stream(24, 0.0, 20260101)
    with serving.Server(200e9, chunk_ms=25.0) as srv:
        mix = serving.serve(srv, reqs, storage="f16", keep_results=True)
    assert mix.completed == len(reqs)
    for k, r in enumerate(reqs):
        with serving.Server(8e9, chunk_ms=25.0, max_active=1) as srv:
            one = serving.serve(srv, [serving.Request(0.0, r.kind, r.edge, r.steps, "f16", r.seed, lid=r.lid,
                                                      tau=r.tau)], keep_results=True)
        np.testing.assert_array_equal(mix.results[k][0], one.results[0][0])
        np.testing.assert_array_equal(mix.results[k][1], one.results[0][1])
    say(f"  {len(reqs)} cases identical")


def _served(cases, **kw):
    """Submit `cases` to one server, drive it to completion, return each
    case's (rho, j) in submission order."""
    with serving.Server(2e9, chunk_ms=0.5, max_chunk_steps=8, **kw) as srv:
        ids = [srv.submit(c) for c in cases]
        for _ in range(10000):
            srv.progress()
            if not srv.busy():
                break
        assert not srv.busy()
        for i in ids:
            assert srv.status(i)["state"] == _capi.CASE_DONE, srv.status(i)
        return [srv.fetch(i) for i in ids]


@pytest.mark.parametrize("storage,dims,variant", [("f32", (30, 22, 18), 0), ("f16", (30, 22, 18), 0),
                                                   ("f16", (160, 20, 18), 11)])
def test_served_channel_equals_plain_channel(storage, dims, variant):
    """A served periodic channel (spheres, seeded perturbation, body force,
    an odd step count) against the channel built outside the server
    (precision.build_channel: periodic comms of the decomposition, ghost
    planes synchronised once, the seam by the copy launch) and stepped by its
    own group: rho and j bit for bit.  The server puts the seam inside the
    step kernel for FP32 and, with packed pairs and rows of 128 cells or
    more, for FP16 (seam wavefronts)."""
    from turbulent_lbm_multigpu_amd import precision

    steps, seed, amp = 37, 424242, 1e-4
    say(f"case: served channel {dims} {storage} variant {variant}, {steps} steps vs build_channel")
    case = serving.make_case("channel", dims, steps, storage, tau=0.51, cs=0.17, body_force=1e-5, fill=0.05,
                             r_min=2.0, r_max=4.0, perturbation=amp, seed=seed)
    (rho_s, j_s), = _served([case], kernel_variant=variant)
    s, g, filled = precision.build_channel(dims, precision.STORAGES[storage], tau=0.51, cs=0.17, gx=1e-5,
                                           sphere_seed=seed, fill=0.05, rmin=2.0, rmax=4.0, amp=amp, seed=seed,
                                           fused_x=0)
    assert filled > 0.02
    g.step(steps)
    g.wait()
    rho, j = s.moments()
    flags = s.store(_capi.FIELD_FLAGS)
    g.close()
    s.close()
    assert (flags == _capi.FLAG_OBSTACLE).any()
    assert np.abs(j).max() > 1e-5   # the flow the check can fail on
    np.testing.assert_array_equal(rho_s, rho.astype(np.float32))
    np.testing.assert_array_equal(j_s, j.astype(np.float32))


@pytest.mark.parametrize("wg,variant", [(64, 0), (128, 8), (192, 3)])
def test_server_small_work_groups(wg, variant):
    """A server created with a work-group size below 256 serves its cases,
    and their results equal the default launch shape's."""
    say(f"case: server wg {wg} variant {variant}")
    cases = lambda: [serving.make_case("cavity", (18, 18, 18), 15, "f32", lid=0.06, seed=5),
                     serving.make_case("channel", (20, 16, 14), 12, "f16", fill=0.03, r_min=2.0, r_max=3.0,
                                       perturbation=1e-4, seed=6)]
    got = _served(cases(), work_group_size=wg, kernel_variant=variant)
    want = _served(cases())
    for (r0, j0), (r1, j1) in zip(got, want):
        np.testing.assert_array_equal(r0, r1)
        np.testing.assert_array_equal(j0, j1)


def test_resident_cell_limit_admits_in_order():
    """max_active_cells: the server never holds more resident cells than the
    limit, except one case larger than the limit on its own (admitted when
    nothing else is resident), and every case's result is the one an
    unlimited server gives."""
    say("case: resident-cell limit")
    cases = lambda: ([serving.make_case("cavity", (24, 24, 24), 14 + k, "f32", seed=k) for k in range(6)]
                     + [serving.make_case("cavity", (40, 40, 40), 9, "f32", seed=9)]
                     + [serving.make_case("channel", (26, 20, 18), 11, "f16", fill=0.03, r_min=2.0, r_max=3.0,
                                          seed=10)])
    limit = 3 * 24 ** 3
    with serving.Server(2e9, chunk_ms=0.5, max_chunk_steps=4, max_active_cells=limit) as srv:
        ids = [srv.submit(c) for c in cases()]
        for _ in range(20000):
            srv.progress()
            if not srv.busy():
                break
        st = srv.stats()
        assert st["completed"] == len(ids) and st["failed"] == 0
        assert 0 < st["cells_peak"] <= max(limit, 40 ** 3)
        got = [srv.fetch(i) for i in ids]
    want = _served(cases())
    for (r0, j0), (r1, j1) in zip(got, want):
        np.testing.assert_array_equal(r0, r1)
        np.testing.assert_array_equal(j0, j1)
