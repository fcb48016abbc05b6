"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Both sides compute with separate IEEE operations in the reference's order
(-ffp-contract=off on both), so FP32 and FP64 results are required to match
bit for bit; the rho/u tolerance the task states (relative 1e-5) is checked as
well, as the weaker contract.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

import turbulent_lbm_multigpu_amd as lbm  # noqa: E402,F401
from turbulent_lbm_multigpu_amd import decomposition as dec  # noqa: E402
from turbulent_lbm_multigpu_amd.solver import LbmGroup, LbmSolver, colocate  # noqa: E402
from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from turbulent_lbm_multigpu_amd import batch  # noqa: E402

RTOL_RHO_U = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_dd(s: LbmSolver):
    return s.store(_capi.FIELD_DD).reshape(19, -1)


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)


def make_pair(dims, bc=(1,) * 6, *, inv_tau, tau=0.0, cs=0.0, lid=0.0, g=(0.0, 0.0, 0.0), forcing=0,
              storage=_capi.STORAGE_F32, wg=256, lid_layer=True, amp=0.0, seed=0, store=True, variant=0):
    dtype = np.float64 if storage == _capi.STORAGE_F64 else np.float32
    o = orc.OracleDomain(dims, bc, inv_tau=inv_tau, tau=tau, smagorinsky_cs=cs, lid_velocity=lid,
                         gravitation=g, forcing=forcing, dtype=dtype)
    s = LbmSolver(dims, bc, inv_tau=inv_tau, tau=tau, smagorinsky_cs=cs, lid_velocity=lid, gravitation=g,
                  forcing=forcing, storage=storage, work_group_size=wg, store_velocity=store,
                  store_density=store, kernel_variant=variant)
    if amp:
        o.reset(amplitude=amp, seed=seed)
        s.init_perturbed(amp, seed)
    if lid_layer:
        origin, size = dec.lid_region(dims)
        o.set_flags(origin, size, orc.FLAG_INJECTION)
        s.set(_capi.FIELD_FLAGS, np.full(size[::-1], orc.FLAG_INJECTION, np.int32), origin, size)
    return o, s


def test_init_matches_oracle():
    say("case: init 17x13x11 perturbed")
    o, s = make_pair((17, 13, 11), inv_tau=1 / 0.6, lid_layer=False, amp=1e-3, seed=20260101)
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    np.testing.assert_array_equal(s.store(_capi.FIELD_FLAGS).ravel(), o.flags)
    s.close()


VARIANTS = [0, 2, 4, 8, 16, 16 | 4]


@pytest.mark.parametrize("steps", [1, 2, 3, 8])
@pytest.mark.parametrize("wg,variant", [(64, 0), (256, 0), (128, 8), (256, 16), (64, 16), (512, 3), (512, 8)])
def test_bgk_cavity_bitwise(steps, wg, variant):
    say(f"case: bgk cavity 19x14x12 steps={steps} wg={wg} variant={variant}")
    o, s = make_pair((19, 14, 12), inv_tau=1 / 0.6, lid=0.05, wg=wg, amp=1e-3, seed=3, variant=variant)
    o.run(steps)
    s.run(steps)
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    np.testing.assert_array_equal(s.store(_capi.FIELD_VELOCITY).reshape(3, -1), o.vel)
    np.testing.assert_array_equal(s.store(_capi.FIELD_DENSITY).ravel(), o.dens)
    s.close()


@pytest.mark.parametrize("steps", [2, 5])
@pytest.mark.parametrize("variant", VARIANTS)
def test_les_bitwise(steps, variant):
    say(f"case: LES 40x36x32 steps={steps} variant={variant}")
    o, s = make_pair((40, 36, 32), inv_tau=1 / 0.5005, tau=0.5005, cs=0.17, lid=0.1, amp=1e-2, seed=20260101,
                     variant=variant)
    o.run(steps)
    s.run(steps)
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    s.close()


@pytest.mark.parametrize("variant", [0, 8, 16])
def test_ghost_faces_and_obstacles_bitwise(variant):
    say(f"case: ghost faces + interior obstacle 36x30x28 variant={variant}")
    o, s = make_pair((36, 30, 28), (8, 8, 1, 8, 8, 1), inv_tau=1 / 0.7, lid=0.04, amp=1e-2, seed=8,
                     variant=variant)
    blk = np.full((2, 3, 4), orc.FLAG_OBSTACLE, np.int32)
    o.set_flags((5, 4, 3), (4, 3, 2), orc.FLAG_OBSTACLE)
    s.set(_capi.FIELD_FLAGS, blk, (5, 4, 3), (4, 3, 2))
    o.run(7)
    s.run(7)
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    np.testing.assert_array_equal(s.store(_capi.FIELD_VELOCITY).reshape(3, -1), o.vel)
    s.close()


@pytest.mark.parametrize("forcing,g", [(0, (1e-4, -2e-4, 3e-5)), (1, (1e-5, 0.0, -2e-6))])
@pytest.mark.parametrize("variant", [0, 16])
def test_forcing_bitwise(forcing, g, variant):
    say(f"case: forcing={forcing} g={g} variant={variant}")
    o, s = make_pair((34, 32, 30), inv_tau=1 / 0.8, lid=0.03, g=g, forcing=forcing, amp=1e-3, seed=1,
                     variant=variant)
    o.run(6)
    s.run(6)
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    np.testing.assert_array_equal(s.store(_capi.FIELD_DENSITY).ravel(), o.dens)
    s.close()


@pytest.mark.parametrize("variant", [0, 8])
def test_fp64_storage_bitwise(variant):
    say(f"case: fp64 LES 32x30x28 variant={variant}")
    o, s = make_pair((32, 30, 28), inv_tau=1 / 0.52, tau=0.52, cs=0.17, lid=0.08, amp=1e-3, seed=4,
                     storage=_capi.STORAGE_F64, variant=variant)
    o.run(9)
    s.run(9)
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    s.close()


def test_config1_cavity_1000_steps():
    """The plain reference method: 64^3 BGK cavity, tau 0.6, lid 0.05, f0 = w,
    1000 steps; rho and u of fluid cells within 1e-5 (and bitwise)."""
    say("case: config 64^3 BGK cavity 1000 steps")
    dims = (64, 64, 64)
    o, s = make_pair(dims, inv_tau=float(np.float32(1.0) / np.float32(0.6)), lid=0.05, store=False)
    o.run(1000)
    s.run(1000)
    s.wait()
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    rho_o, j_o = o.moments()
    rho_g, j_g = s.moments()
    fluid = o.flags == orc.FLAG_FLUID
    assert rel_err(rho_g.ravel()[fluid], rho_o[fluid]) <= RTOL_RHO_U
    assert rel_err(j_g.reshape(3, -1)[:, fluid], j_o[:, fluid]) <= RTOL_RHO_U
    # the flow is driven by the lid: the cells under the lid move in +x
    assert j_o[0].reshape(64, 64, 64)[32, 61, 32] > 0.01
    s.close()


def test_regions_norm_and_checksum():
    say("case: regions / norm filter / checksum")
    dims = (12, 10, 9)
    o, s = make_pair(dims, inv_tau=1 / 0.6, lid=0.05, amp=1e-3, seed=2)
    o.run(3)
    s.run(3)
    origin, size = (2, 3, 1), (5, 4, 6)
    reg = s.store(_capi.FIELD_DD, origin, size).reshape(19, -1)
    np.testing.assert_array_equal(reg, o.dd_region(origin, size))
    rng = np.random.default_rng(1)
    src = rng.random((19, int(np.prod(size)))).astype(np.float32)
    s.set(_capi.FIELD_DD, src, origin, size, norm=(1, 0, 0))
    o.set_dd_region(origin, size, src, norm=(1, 0, 0))
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    s.set(_capi.FIELD_DD, src, origin, size)
    o.set_dd_region(origin, size, src)
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    vel = s.store(_capi.FIELD_VELOCITY).reshape(3, -1)
    fl = s.store(_capi.FIELD_FLAGS).ravel()
    ref = np.float32(0)
    for a in range(vel.shape[1]):   # src/CLbmSolver.hpp:1114-1117, float accumulation
        if fl[a] == orc.FLAG_FLUID:
            ref = np.float32(ref + np.float32(np.float32(vel[0, a] + vel[1, a]) + vel[2, a]))
    assert s.velocity_checksum() == ref
    with pytest.raises(_capi.LbmError) as ei:
        s.store(_capi.FIELD_DD, (10, 0, 0), (5, 1, 1))
    assert ei.value.code == _capi.LBM_ERR_OUT_OF_RANGE
    s.close()


def test_mass_energy_fp64_reduction():
    say("case: FP64 mass/energy reduction")
    o, s = make_pair((20, 18, 16), inv_tau=1 / 0.6, lid=0.05, amp=1e-2, seed=6)
    o.run(5)
    s.run(5)
    m, k = s.mass_energy()
    rho, j = o.moments()
    fluid = o.flags == orc.FLAG_FLUID
    m_ref = rho[fluid].sum()
    k_ref = (0.5 * (j[:, fluid] ** 2).sum(0) / rho[fluid]).sum()
    assert abs(m - m_ref) <= 1e-12 * m_ref
    assert abs(k - k_ref) <= 1e-10 * k_ref
    s.close()


def build_gpu_group(domain_size, nums, *, lid, amp, seed, storage=_capi.STORAGE_F32, mode=0, graph=True,
                    periodic=(False, False, False), colocated=False, **kw):
    mono = dec.monolithic_size(domain_size, nums)
    subs = dec.decompose(domain_size, nums, periodic=periodic)
    solvers = [LbmSolver(sd.domain.size, [sd.bc[0][0], sd.bc[0][1], sd.bc[1][0], sd.bc[1][1], sd.bc[2][0],
                                          sd.bc[2][1]], lid_velocity=lid, storage=storage, store_velocity=True,
                         store_density=True, **kw) for sd in subs]
    if colocated:
        colocate(solvers)
    for sd, s in zip(subs, solvers):
        s.init_perturbed(amp, seed, 0, dec.key_origin(sd), mono)
        if sd.has_lid:
            o, sz = dec.lid_region(sd.domain.size)
            s.set(_capi.FIELD_FLAGS, np.full(sz[::-1], orc.FLAG_INJECTION, np.int32), o, sz)
    grp = LbmGroup(solvers)
    for sd in subs:
        for c in sd.comms:
            grp.add_comm(sd.uid, c.dst_id, c.send_origin, c.recv_origin, c.send_size, c.comm_direction)
    grp.set_exchange_mode(mode)
    grp.set_graph(graph)
    return grp, subs, mono


@pytest.mark.parametrize("graph,colocated", [(True, False), (False, False), (True, True)])
def test_group_2x2x2_matches_oracle_decomposition(graph, colocated):
    """Also with the eight subdomains moved into one allocation
    (lbm_solver_colocate): distributions, velocity and density outputs."""
    say(f"case: group 2x2x2 of 14^3 graph={graph} colocated={colocated}")
    from tests.test_oracle import build_oracle_decomposition
    domain_size, nums = (28, 28, 28), (2, 2, 2)
    kw = dict(inv_tau=1 / 0.5005, tau=0.5005, smagorinsky_cs=0.17)
    multi, mono = build_oracle_decomposition(domain_size, nums, lid=0.1, amp=1e-4, seed=20260101, **kw)
    grp, subs, _ = build_gpu_group(domain_size, nums, lid=0.1, amp=1e-4, seed=20260101, graph=graph,
                                   colocated=colocated, **kw)
    for steps in (3, 6):   # odd then even totals: both sync kinds, graph replays in between
        multi.run(steps)
        grp.step(steps)
        grp.wait()
        for sub, s in zip(multi.subs, grp.solvers):
            np.testing.assert_array_equal(gpu_dd(s), sub["domain"].dd)
            if colocated:
                np.testing.assert_array_equal(s.store(_capi.FIELD_VELOCITY).ravel(), sub["domain"].vel.ravel())
                np.testing.assert_array_equal(s.store(_capi.FIELD_DENSITY).ravel(), sub["domain"].dens.ravel())
    grp.close()
    for s in grp.solvers[::-1]:   # the shared allocation goes with the last solver
        s.close()


def test_colocate_keeps_state_and_rejects_group_members():
    """Colocation moves buffers with their contents (distributions, flags,
    outputs); a solver already colocated or in a group is refused."""
    say("case: colocate state and errors")
    from turbulent_lbm_multigpu_amd._capi import LbmError
    solvers = [LbmSolver(d, (1,) * 6, inv_tau=1 / 0.55, tau=0.55, smagorinsky_cs=0.17, lid_velocity=0.06,
                         store_velocity=v, store_density=v) for d, v in
               (((21, 17, 13), True), ((16, 16, 16), False), ((30, 9, 11), True))]
    before = []
    for k, s in enumerate(solvers):
        s.init_perturbed(1e-3, 40 + k)
        s.run(3 + k)
        s.wait()
        before.append((gpu_dd(s), s.store(_capi.FIELD_FLAGS), s.store(_capi.FIELD_VELOCITY) if k != 1 else None))
    colocate(solvers)
    for s, (dd, fl, vel) in zip(solvers, before):
        np.testing.assert_array_equal(gpu_dd(s), dd)
        np.testing.assert_array_equal(s.store(_capi.FIELD_FLAGS), fl)
        if vel is not None:
            np.testing.assert_array_equal(s.store(_capi.FIELD_VELOCITY), vel)
    with pytest.raises(LbmError) as e:
        colocate(solvers[:1])
    assert e.value.code == _capi.LBM_ERR_STATE
    extra = LbmSolver((12, 12, 12), (1,) * 6, inv_tau=1 / 0.6)
    grp = LbmGroup([extra])
    with pytest.raises(LbmError) as e:
        colocate([extra])
    assert e.value.code == _capi.LBM_ERR_STATE
    grp.close()
    extra.close()
    # stepping after the move, then destroy in creation order
    for s in solvers:
        s.run(2)
        s.wait()
        assert np.isfinite(gpu_dd(s)).all()
    for s in solvers:
        s.close()


def test_group_decomposed_equals_monolithic_and_minimal_exchange():
    """The reference's validation mode on the GPU (src/main.cpp:309-408), plus
    the minimal exchange (5 populations after alpha) giving the same interiors."""
    say("case: group 3x1x2 vs monolithic, exchange modes 0/1")
    domain_size, nums = (36, 16, 24), (3, 1, 2)
    kw = dict(inv_tau=1 / 0.55, tau=0.55, smagorinsky_cs=0.17)
    g0, subs, mono = build_gpu_group(domain_size, nums, lid=0.07, amp=1e-3, seed=9, mode=0, **kw)
    g1, _, _ = build_gpu_group(domain_size, nums, lid=0.07, amp=1e-3, seed=9, mode=1, **kw)
    one = LbmSolver(mono, (1,) * 6, lid_velocity=0.07, store_velocity=True, **kw)
    one.init_perturbed(1e-3, 9)
    o, sz = dec.lid_region(mono)
    one.set(_capi.FIELD_FLAGS, np.full(sz[::-1], orc.FLAG_INJECTION, np.int32), o, sz)
    steps = 21
    g0.step(steps)
    g1.step(steps)
    one.run(steps)
    ref = one.store(_capi.FIELD_DD)
    for sd, s0, s1 in zip(subs, g0.solvers, g1.solvers):
        s3 = sd.domain.size
        ko = dec.key_origin(sd)
        a = s0.store(_capi.FIELD_DD)[:, 1:-1, 1:-1, 1:-1]
        b = s1.store(_capi.FIELD_DD)[:, 1:-1, 1:-1, 1:-1]
        r = ref[:, ko[2] + 1:ko[2] + s3[2] - 1, ko[1] + 1:ko[1] + s3[1] - 1, ko[0] + 1:ko[0] + s3[0] - 1]
        np.testing.assert_array_equal(a, r)
        np.testing.assert_array_equal(b, r)
    for g in (g0, g1):
        g.close()
        for s in g.solvers:
            s.close()
    one.close()


@pytest.mark.parametrize("alpha_wg", ["default", 0])
def test_full_size_benchmark_decomposition_equals_monolithic(alpha_wg):
    """The benchmark workload at its full size, launched as bench.py launches
    it by default (256^3 LES as 2x2x2 subdomains of 128^3, kernel variant 3,
    x faces in the kernel, exchange mode 1, one copy launch per half-step,
    graph replay, alpha at its own work-group size), and with one size for
    both kernels, against a single 254^3 domain: the reference's validation
    invariant (src/main.cpp:358-405), interior rho and j bit-identical after
    10 steps."""
    import bench

    a = bench.arg_parser().parse_args([])
    wa = a.wg_alpha if alpha_wg == "default" else alpha_wg
    say(f"case: full-size benchmark group (wg {a.wg}, alpha {wa}) vs monolithic 254^3, 10 steps")
    w = bench.WORKLOAD
    grp = bench.build_group(exchange_mode=a.exchange_mode, graph=True, wg=a.wg, kernel_variant=a.kernel_variant,
                            fused_x=bool(a.fused_x), wg_alpha=wa)
    mono = dec.monolithic_size(w["domain"], w["nums"])
    one = LbmSolver(mono, (1,) * 6, inv_tau=1.0 / w["tau"], tau=w["tau"], smagorinsky_cs=w["cs"],
                    lid_velocity=w["lid"], kernel_variant=3)
    one.init_perturbed(w["amp"], w["seed"])
    o, sz = dec.lid_region(mono)
    one.set(_capi.FIELD_FLAGS, np.full(sz[::-1], orc.FLAG_INJECTION, np.int32), o, sz)
    grp.step(10)
    one.run(10)
    grp.wait()
    one.wait()
    for sd, s in zip(dec.decompose(w["domain"], w["nums"]), grp.solvers):
        s3, ko = sd.domain.size, dec.key_origin(sd)
        size = tuple(s3[a] - 2 for a in range(3))
        rho_a, j_a = s.moments((1, 1, 1), size)
        rho_b, j_b = one.moments(tuple(ko[a] + 1 for a in range(3)), size)
        np.testing.assert_array_equal(rho_a, rho_b)
        np.testing.assert_array_equal(j_a, j_b)
    grp.close()
    for s in grp.solvers:
        s.close()
    one.close()


@pytest.mark.parametrize("storage", [_capi.STORAGE_F32, _capi.STORAGE_F16_SHIFTED])
def test_full_size_precision_channel_fused_equals_copies(storage):
    """The precision workload at its full size (384x192x192 periodic LES
    channel with spheres and a body force): x faces in the kernel by run
    tiles (a 384-cell row is wider than a work-group) equal the copy
    exchange, rho and j bit-identical after 6 steps."""
    say(f"case: full-size channel 384x192x192 storage {storage}, fused vs copies")
    from turbulent_lbm_multigpu_amd import precision

    out = []
    for fused in (True, False):
        s, g, _ = precision.build_channel((384, 192, 192), storage, kernel_variant=3, fused_x=fused)
        g.step(6)
        g.wait()
        out.append(s.moments())
        g.close()
        s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("storage", [_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
def test_full_size_precision_channel_packed_pairs(storage):
    """The precision workload at its full size with two cells per lane as
    packed 32-bit words and the x faces by the copy launch, against one cell
    per lane with the x faces in the kernel, and against packed pairs with
    the x faces in the kernel: every stored bit identical after 7 steps."""
    say(f"case: full-size channel 384x192x192 storage {storage}, packed pairs vs one cell per lane")
    from turbulent_lbm_multigpu_amd import precision

    out = []
    for variant, fused in ((11, 0), (3, 1), (11, 1)):
        s, g, _ = precision.build_channel((384, 192, 192), storage, kernel_variant=variant, fused_x=fused)
        g.step(7)
        g.wait()
        out.append(s.store(_capi.FIELD_DD_RAW))
        g.close()
        s.close()
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


@pytest.mark.parametrize("storage", [_capi.STORAGE_F32, _capi.STORAGE_F64, _capi.STORAGE_F16_SHIFTED,
                                     _capi.STORAGE_BF16_SHIFTED])
@pytest.mark.parametrize("dims,variant", [((40, 18, 16), 3), ((160, 12, 10), 11)])
def test_fused_x_auto_picks_per_storage(storage, dims, variant):
    """LBM_FUSED_X_AUTO: the periodic seam of a channel goes inside the step
    kernel for FP32/FP64 (no exchange launch); for FP16/BF16 it goes inside
    the kernel as seam wavefronts when the launch packs pairs and the rows
    are at least 128 cells long, else through the copy launch (one per
    step); the fields equal both explicit modes bit for bit. Other mode
    values are rejected."""
    say(f"case: fused x auto, channel {dims} storage {storage} variant {variant}")
    from turbulent_lbm_multigpu_amd import precision

    out, launches = [], []
    for mode in (_capi.FUSED_X_AUTO, 1, 0):
        s, g, _ = precision.build_channel(dims, storage, tau=0.55, gx=2e-5, fill=0.04, rmin=2.0, rmax=3.0,
                                          kernel_variant=variant, fused_x=mode)
        g.set_timing(True)
        g.step(8)
        g.wait()
        launches.append(g.timing()["exchange_launches"])
        out.append(s.store(_capi.FIELD_DD_RAW))
        if mode == 0:
            with pytest.raises(RuntimeError):
                g.set_fused_x(3)
        g.close()
        s.close()
    two_byte = storage in (_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED)
    seam = two_byte and variant == 11 and dims[0] >= 128
    # one launch per step (a step is one half-step kernel, alpha or beta)
    assert launches[1] == 0 and launches[2] == 8
    assert launches[0] == (8 if two_byte and not seam else 0)
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


@pytest.mark.parametrize("storage", [_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
@pytest.mark.parametrize("kind", ["bgk", "les", "les_body_force", "les_reference_forcing"])
@pytest.mark.parametrize("variant", [0, 11])
def test_two_byte_half_steps_equal_oracle_to_storage_rounding(storage, kind, variant):
    """The two-byte storages collide in their own form (pair sums and
    differences, the hardware square root and reciprocal in the LES rate),
    so they are held to the exact update rather than to FP32's rounding
    order: from the same state, exactly representable in the storage, one
    beta and then one alpha half-step must give every slot of the FP64
    oracle to within half a storage ulp of f - w_i (the output rounding)
    + 1e-7 (FP32 arithmetic on f - w_i), and the velocity and density
    outputs (FP32, no storage rounding) to within 2e-7.  (The FP32 oracle
    itself is up to 2.4e-7 from FP64 here: it collides the full f, whose
    LES stress cancels the w terms in FP32.)  A dropped term, a wrong sign or
    index in the pair algebra, or a wrong LES constant moves a slot by 1e-5
    or more at these amplitudes (FP16's half ulp at |f - w| = 0.03 is
    7.6e-6; tests/test_two_byte_tolerance_cpu.py plants such mistakes in the
    oracle and checks that they exceed this tolerance)."""
    from tests.test_two_byte_tolerance_cpu import W19, oracle_domain, tolerance

    cfg = {"bgk": dict(tau=0.6, cs=0.0, forcing=1, g=(0.0, 0.0, 0.0)),
           "les": dict(tau=0.505, cs=0.17, forcing=1, g=(0.0, 0.0, 0.0)),
           "les_body_force": dict(tau=0.505, cs=0.17, forcing=1, g=(2e-4, -1e-4, 1.5e-4)),
           "les_reference_forcing": dict(tau=0.505, cs=0.17, forcing=0, g=(2e-4, -1e-4, 1.5e-4))}[kind]
    say(f"case: two-byte half-steps vs oracle, storage {storage} {kind} variant {variant}")
    fmt = "f16" if storage == _capi.STORAGE_F16_SHIFTED else "bf16"
    _, s = make_pair((20, 16, 14), inv_tau=1 / cfg["tau"], tau=cfg["tau"], cs=cfg["cs"], lid=0.08, g=cfg["g"],
                     forcing=cfg["forcing"], storage=storage, amp=2e-2, seed=11, variant=variant)
    o = oracle_domain(np.float64, **cfg)
    o.dd[...] = gpu_dd(s)   # the oracle starts from the device's rounded state
    for half in ("beta", "alpha"):
        o.step()
        s.run(1)
        got, want = gpu_dd(s).astype(np.float64), o.dd
        tol = tolerance(got, want, fmt)
        err = np.abs(got - want)
        say(f"  {half}: max |slot error| {err.max():.3e}, max error / tolerance {(err / tol).max():.3f}")
        assert (err <= tol).all(), f"{half}: {np.count_nonzero(err > tol)} slots beyond the storage rounding"
        np.testing.assert_allclose(s.store(_capi.FIELD_VELOCITY).reshape(3, -1), o.vel, rtol=0, atol=2e-7)
        np.testing.assert_allclose(s.store(_capi.FIELD_DENSITY).ravel(), o.dens, rtol=0, atol=2e-7)
        o.dd[...] = got   # continue from the device's rounded state
    s.close()


@pytest.mark.parametrize("storage", [_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
def test_shifted_storage_error_vs_fp64(storage):
    """FP16/BF16 storage of f - w_i with FP32 collision stays close to FP64."""
    say(f"case: shifted storage {storage} vs fp64")
    dims = (24, 24, 24)
    kw = dict(inv_tau=1 / 0.55, tau=0.55, smagorinsky_cs=0.17, lid_velocity=0.1)
    ref = LbmSolver(dims, storage=_capi.STORAGE_F64, **kw)
    low = LbmSolver(dims, storage=storage, **kw)
    for s in (ref, low):
        s.init_perturbed(1e-3, 5)
        o, sz = dec.lid_region(dims)
        s.set(_capi.FIELD_FLAGS, np.full(sz[::-1], orc.FLAG_INJECTION, np.int32), o, sz)
        s.run(200)
    r_rho, r_j = ref.moments()
    l_rho, l_j = low.moments()
    fluid = ref.store(_capi.FIELD_FLAGS) == orc.FLAG_FLUID
    e_rho = np.abs(l_rho[fluid] - r_rho[fluid]).max()
    e_u = np.abs(l_j[:, fluid] - r_j[:, fluid]).max() / 0.1
    say(f"  max |drho| = {e_rho:.3e}, max |du|/U = {e_u:.3e}")
    bound_rho, bound_u = (2e-3, 2e-2) if storage == _capi.STORAGE_F16_SHIFTED else (2e-2, 2e-1)
    assert e_rho < bound_rho
    assert e_u < bound_u
    ref.close()
    low.close()


def fluid_slots(s: LbmSolver):
    """The 19 slots of every FLUID cell (what the next half-step reads)."""
    fl = s.store(_capi.FIELD_FLAGS).reshape(-1)
    return gpu_dd(s).reshape(19, -1)[:, fl == orc.FLAG_FLUID]


def fused_vs_copy(domain_size, nums, mode, *, wg=256, periodic=(False, False, False), stops=(), step_runs=(5, 8),
                  storage=_capi.STORAGE_F32, variant=0, alpha_wg=0, fuse=True):
    """Two identical groups, one with x faces inside the step kernel (and,
    with alpha_wg, alpha launched at that work-group size): every slot of
    every subdomain is bit-identical after each run of steps."""
    kw = dict(inv_tau=1 / 0.52, tau=0.52, smagorinsky_cs=0.17, work_group_size=wg, storage=storage,
              kernel_variant=variant)
    ref, _, _ = build_gpu_group(domain_size, nums, lid=0.08, amp=1e-3, seed=11, mode=mode, periodic=periodic, **kw)
    fus, _, _ = build_gpu_group(domain_size, nums, lid=0.08, amp=1e-3, seed=11, mode=mode, periodic=periodic, **kw)
    fus.set_fused_x(True if fuse else 0)
    if alpha_wg:
        fus.set_alpha_work_group(alpha_wg)
        assert fus.alpha_work_group() == alpha_wg and fus.launch()[0] == wg
    for member, at in stops:
        ref.set_member_stop(member, at)
        fus.set_member_stop(member, at)
    for steps in step_runs:
        ref.step(steps)
        fus.step(steps)
        ref.wait()
        fus.wait()
        for a, b in zip(ref.solvers, fus.solvers):
            np.testing.assert_array_equal(gpu_dd(a), gpu_dd(b))
    for g in (ref, fus):
        g.close()
        for s in g.solvers:
            s.close()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("domain_size,nums", [((28, 28, 28), (2, 2, 2)), ((36, 16, 24), (3, 1, 2)),
                                              ((56, 12, 10), (4, 1, 1))])
def test_fused_x_faces_match_copy_exchange(mode, domain_size, nums):
    """x faces exchanged inside the step kernel (chains of 2, 3 and 4
    subdomains along x): all slots bit-identical to the copy-based exchange,
    after odd and even step counts."""
    say(f"case: fused x faces {domain_size} {nums} mode {mode}")
    fused_vs_copy(domain_size, nums, mode)


@pytest.mark.parametrize("wg", [64, 128, 256, 512])
@pytest.mark.parametrize("domain_size", [(56, 12, 10), (200, 12, 10), (260, 12, 10)])
def test_fused_x_chain_fits_work_group_or_copies(wg, domain_size):
    """A chain tile holds whole rows of every member, or for wider rows the
    64-cell runs next to the faces (the interior runs in tiles of their own);
    chains whose face runs do not fit the work-group keep the copy exchange.
    Results never depend on which."""
    nums = (4, 1, 1) if domain_size[0] == 56 else (2, 1, 1)
    say(f"case: fused x chain {domain_size} {nums} wg {wg}")
    fused_vs_copy(domain_size, nums, 1, wg=wg, step_runs=(3, 4))


@pytest.mark.parametrize("domain_size,nums,periodic,wg,alpha_wg,fuse,stops", [
    ((256, 12, 10), (2, 1, 1), False, 256, 512, True, ()),    # 128-cell rows: alpha's tiles hold 2 rows per member
    ((256, 12, 10), (2, 1, 1), False, 512, 256, True, ()),    # ... and the reverse
    ((256, 12, 10), (2, 1, 1), True, 256, 512, True, ()),     # a periodic ring of two
    ((256, 12, 10), (2, 1, 1), False, 256, 512, True, ((1, 5),)),  # a frozen member
    ((56, 12, 10), (4, 1, 1), False, 64, 512, True, ()),      # short rows, many per tile
    ((28, 28, 28), (2, 2, 2), False, 256, 128, True, ()),     # x, y and z faces
    ((28, 28, 28), (2, 2, 2), False, 256, 512, False, ()),    # copy exchange, alpha's own tile table
    ((520, 12, 10), (2, 1, 1), False, 512, 64, True, ()),     # the plan is made for the smaller size
])
def test_alpha_work_group_of_its_own(domain_size, nums, periodic, wg, alpha_wg, fuse, stops):
    """Alpha launched at a work-group size other than beta's: all slots
    equal to one work-group size and the copy exchange."""
    say(f"case: alpha wg {alpha_wg} beta wg {wg} {domain_size} {nums} periodic {periodic} fuse {fuse} stops {stops}")
    fused_vs_copy(domain_size, nums, 1, wg=wg, alpha_wg=alpha_wg, fuse=fuse, stops=stops,
                  periodic=(periodic, False, False), step_runs=(3, 4))


@pytest.mark.parametrize("storage", [_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
def test_alpha_work_group_of_its_own_seam_pairs(storage):
    """Packed pairs with the periodic seam in one wavefront per row, alpha
    at 512 and beta at 256: all slots equal to the copy exchange."""
    say(f"case: alpha wg 512 seam pairs storage {storage}")
    fused_vs_copy((384, 10, 8), (1, 1, 1), 1, wg=256, alpha_wg=512, periodic=(True, False, False), storage=storage,
                  variant=11, step_runs=(3, 4))
    fused_vs_copy((384, 10, 8), (1, 1, 1), 1, wg=128, alpha_wg=512, periodic=(True, False, False), storage=storage,
                  variant=11, step_runs=(3, 4))


@pytest.mark.parametrize("storage", [_capi.STORAGE_F64, _capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
@pytest.mark.parametrize("variant", [0, 3])
def test_fused_x_storages_and_cache_policy(storage, variant):
    """Chain tiles in every storage type, with the default and the
    non-temporal cache policy: all slots equal to the copy exchange."""
    say(f"case: fused x 2x2x2 storage {storage} variant {variant}")
    fused_vs_copy((32, 28, 24), (2, 2, 2), 1, storage=storage, variant=variant, step_runs=(3, 4))


@pytest.mark.parametrize("storage", [_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
@pytest.mark.parametrize("domain_size,nums,periodic,wg", [
    ((384, 10, 8), (1, 1, 1), True, 256),     # ring of one: each row's seam in one wavefront (3 per row)
    ((384, 10, 8), (1, 1, 1), True, 128),     # ... two wavefronts per work-group
    ((384, 10, 8), (1, 1, 1), True, 512),     # ... eight: a work-group ends inside a row
    ((640, 10, 8), (1, 1, 1), True, 64),      # 5 wavefronts per row, one per work-group
    ((128, 12, 10), (1, 1, 1), True, 256),    # the shortest seam row: one wavefront per row
    ((200, 12, 10), (1, 1, 1), True, 192),    # a row of 100 pairs: the second wavefront half full
    ((28, 12, 10), (1, 1, 1), True, 256),     # rows under 128 cells: one-cell chain tiles
    ((56, 12, 10), (2, 1, 1), False, 256),    # a line of two: one-cell chain tiles
    ((39, 12, 10), (3, 1, 1), False, 256),    # odd rows
])
def test_fused_x_packed_pairs(storage, domain_size, nums, periodic, wg):
    """Two-byte storages with two cells per work-item (variants 8 and 11)
    and the x faces inside the kernel: seam wavefronts for periodic rings of
    one, one-cell chain tiles otherwise; every slot equal to the copy
    exchange after odd and even step counts."""
    for variant in (11, 8):
        say(f"case: fused x packed pairs {domain_size} {nums} periodic {periodic} wg {wg} storage {storage} "
            f"variant {variant}")
        fused_vs_copy(domain_size, nums, 1, wg=wg, periodic=(periodic, False, False), storage=storage,
                      variant=variant, step_runs=(3, 4))


@pytest.mark.parametrize("domain_size,nums", [((42, 12, 10), (3, 1, 1)), ((768, 10, 8), (2, 1, 1))])
def test_fused_x_packed_pairs_with_frozen_member(domain_size, nums):
    """A frozen member inside a chain of a two-byte storage launched with
    packed pairs (whole-row and run tiles): its slots as they stand feed its
    partner, as the copies do."""
    say(f"case: fused x packed pairs {domain_size} {nums} with a frozen member")
    stops = ((1, 3), (2, 6)) if nums[0] == 3 else ((1, 3),)
    fused_vs_copy(domain_size, nums, 1, stops=stops, step_runs=(5, 6), storage=_capi.STORAGE_F16_SHIFTED,
                  variant=11)


def test_seam_rings_in_a_batch_with_a_frozen_member():
    """Seam rings of one in a batch: two periodic channels of packed pairs
    (FP16), one frozen after 3 half-steps, next to a cavity, stepped in one
    group: each equals the same case stepped alone with copies."""
    say("case: seam rings in a batch with a frozen member")
    from turbulent_lbm_multigpu_amd import precision

    def channel(fused):
        return precision.build_channel((160, 12, 10), _capi.STORAGE_F16_SHIFTED, tau=0.55, gx=2e-5, fill=0.04,
                                       rmin=2.0, rmax=3.0, amp=1e-3, seed=5, kernel_variant=11, fused_x=fused)

    solo = []
    for stop in (None, 3):
        s, g, _ = channel(0)
        g.step(3 if stop else 8)
        g.wait()
        solo.append(s.store(_capi.FIELD_DD_RAW))
        g.close()
        s.close()
    grp = LbmGroup.empty(_capi.STORAGE_F16_SHIFTED, 256, 11)
    members = []
    for stop in (None, 3):
        s, g, _ = channel(0)   # set up with its own group (ghost planes synchronised), then moved
        g.close()
        idx = grp.add_member(s)
        grp.add_comm(idx, idx, (1, 0, 0), (0, 0, 0), (1, 12, 10), (1, 0, 0))
        grp.add_comm(idx, idx, (158, 0, 0), (159, 0, 0), (1, 12, 10), (-1, 0, 0))
        if stop:
            grp.set_member_stop(idx, stop)
        members.append(s)
    cav = batch.build_solver(batch.Cavity((20, 18, 16), 0.07, 0.53), storage=_capi.STORAGE_F16_SHIFTED,
                             kernel_variant=11)
    grp.add_member(cav)
    grp.set_exchange_mode(1)
    grp.set_fused_x(_capi.FUSED_X_AUTO)
    grp.set_timing(True)
    grp.step(8)
    grp.wait()
    assert grp.timing()["exchange_launches"] == 0   # both seams inside the kernel
    for s, want in zip(members, solo):
        np.testing.assert_array_equal(s.store(_capi.FIELD_DD_RAW), want)
    grp.close()
    for s in members + [cav]:
        s.close()


def test_fused_x_periodic_ring_of_two():
    """x periodic across two subdomains: a ring chain (each member feeds
    both faces of the other)."""
    say("case: fused x periodic ring (2,1,1)")
    fused_vs_copy((36, 14, 12), (2, 1, 1), 1, periodic=(True, False, False), step_runs=(5, 6))


@pytest.mark.parametrize("domain_size,nums", [((42, 12, 10), (3, 1, 1)), ((260, 10, 8), (2, 1, 1))])
def test_fused_x_with_frozen_member(domain_size, nums):
    """A member frozen at its own step count inside a chain (whole-row and
    one-row tiles): the frozen sender's slots as they stand feed its partner,
    as the copies do."""
    say(f"case: fused x chain {domain_size} {nums} with a frozen member")
    stops = ((1, 3), (2, 6)) if nums[0] == 3 else ((1, 3),)
    fused_vs_copy(domain_size, nums, 1, stops=stops, step_runs=(5, 6))


def test_fused_x_toggled_between_graph_replays():
    """x faces switched into the kernel and back out between steps of a group
    that replays a captured graph (the graph is captured again whenever a
    table it holds moves): all slots equal to a group that always copies."""
    say("case: fused x toggled between graph replays")
    kw = dict(inv_tau=1 / 0.52, tau=0.52, smagorinsky_cs=0.17)
    ref, _, _ = build_gpu_group((32, 28, 24), (2, 2, 2), lid=0.08, amp=1e-3, seed=11, mode=1, **kw)
    tog, _, _ = build_gpu_group((32, 28, 24), (2, 2, 2), lid=0.08, amp=1e-3, seed=11, mode=1, **kw)
    for steps, fused in ((4, False), (5, True), (3, False), (6, True)):
        tog.set_fused_x(fused)
        ref.step(steps)
        tog.step(steps)
        ref.wait()
        tog.wait()
        for a, b in zip(ref.solvers, tog.solvers):
            np.testing.assert_array_equal(gpu_dd(a), gpu_dd(b))
    for g in (ref, tog):
        g.close()
        for s in g.solvers:
            s.close()


def test_group_2x2x2_fused_matches_oracle_decomposition():
    """The fused x faces against the oracle's copy-based decomposition (all
    slots, exchange of all 19 populations)."""
    say("case: fused group 2x2x2 of 14^3 vs oracle")
    from tests.test_oracle import build_oracle_decomposition
    domain_size, nums = (28, 28, 28), (2, 2, 2)
    kw = dict(inv_tau=1 / 0.5005, tau=0.5005, smagorinsky_cs=0.17)
    multi, _ = build_oracle_decomposition(domain_size, nums, lid=0.1, amp=1e-4, seed=20260101, **kw)
    grp, _, _ = build_gpu_group(domain_size, nums, lid=0.1, amp=1e-4, seed=20260101, **kw)
    grp.set_fused_x(True)
    for steps in (3, 6):
        multi.run(steps)
        grp.step(steps)
        grp.wait()
        for sub, s in zip(multi.subs, grp.solvers):
            np.testing.assert_array_equal(gpu_dd(s), sub["domain"].dd)
    grp.close()
    for s in grp.solvers:
        s.close()


@pytest.mark.parametrize("dims", [(24, 18, 16), (300, 9, 7), (321, 7, 6)])
def test_fused_x_periodic_channel(dims):
    """A periodic channel (self comms across x) with fused x faces, a ring of
    one (whole rows, and two-row run tiles with interior tiles when X = 300, the last row tile half full):
    FLUID slots equal to the copy-based exchange after 17 steps."""
    say(f"case: fused x periodic channel {dims}")
    from turbulent_lbm_multigpu_amd import precision

    out = []
    for fused in (False, True):
        s, g, _ = precision.build_channel(dims, _capi.STORAGE_F32, tau=0.55, gx=2e-5, fill=0.04, rmin=2.0,
                                          rmax=3.0, amp=1e-3, seed=9)
        g.set_fused_x(fused)
        g.step(17)
        g.wait()
        out.append(fluid_slots(s))
        g.close()
        s.close()
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("storage", [_capi.STORAGE_F16_SHIFTED, _capi.STORAGE_BF16_SHIFTED])
@pytest.mark.parametrize("dims", [(20, 18, 16), (21, 17, 19), (64, 12, 10)])
def test_two_byte_storage_launch_shapes_identical(storage, dims):
    """FP16/BF16 storage: 1, 2 and 4 cells per lane (packed 32-bit words at
    any 2-byte alignment, element accesses where a pair straddles a layout
    block), with and without the non-temporal policy, work-groups
    64/128/256, give bit-identical states."""
    say(f"case: two-byte storage {storage} dims {dims} launch shapes")
    out = []
    for variant, wg in ((0, 256), (3, 256), (8, 256), (8, 64), (11, 256), (11, 128), (16, 128), (11, 512), (3, 512)):
        s = LbmSolver(dims, (1,) * 6, inv_tau=1 / 0.53, tau=0.53, smagorinsky_cs=0.17, lid_velocity=0.06,
                      storage=storage, kernel_variant=variant, work_group_size=wg)
        s.init_perturbed(1e-3, 17)
        o, sz = dec.lid_region(dims)
        s.set(_capi.FIELD_FLAGS, np.full(sz[::-1], orc.FLAG_INJECTION, np.int32), o, sz)
        s.run(9)
        out.append(gpu_dd(s))
        s.close()
    for other in out[1:]:
        np.testing.assert_array_equal(out[0], other)


def test_two_byte_channel_pairs_with_fused_x():
    """Vector lanes next to fused periodic x faces (edge lanes take the
    element path): FP16 channel, 2 cells per lane == 1 cell per lane, with
    the x faces in the kernel or by the copy launch."""
    say("case: fp16 channel pairs + fused x")
    from turbulent_lbm_multigpu_amd import precision

    out = []
    for variant, fused in ((0, True), (8, True), (11, True), (11, False)):
        s, g, _ = precision.build_channel((32, 18, 16), _capi.STORAGE_F16_SHIFTED, tau=0.55, gx=2e-5, fill=0.04,
                                          rmin=2.0, rmax=3.0, amp=1e-3, seed=9, kernel_variant=variant)
        g.set_fused_x(fused)
        g.step(15)
        g.wait()
        out.append(gpu_dd(s))
        g.close()
        s.close()
    for other in out[1:]:
        np.testing.assert_array_equal(out[0], other)
