"""Seeded random single domains and decompositions against the CPU oracle,
bit for bit.

Each case draws a domain size, the six face flags (obstacle, fluid, velocity
injection or ghost: the linear index wrap of the reference then routes the
boundary populations, and the oracle does the same), an interior obstacle
box, the lid layer, BGK or Smagorinsky LES with a random tau and Cs, no
forcing, the reference's forcing or the body force, FP32 or FP64 storage, a
random launch shape and an odd or even step count.  Distributions, velocity
and density must equal the oracle's exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as orc  # noqa: E402
from turbulent_lbm_multigpu_amd import _capi  # noqa: E402
from tests.test_gpu_parity import gpu_dd, make_pair, say  # noqa: E402

FACES = [orc.FLAG_OBSTACLE, orc.FLAG_FLUID, orc.FLAG_INJECTION, orc.FLAG_GHOST]
VARIANTS = [0, 1, 2, 3, 8, 11, 16]


def draw(case: int):
    rng = np.random.default_rng(7_000 + case)
    dims = [int(rng.integers(6, 34)) for _ in range(3)]
    if rng.random() < 0.2:
        dims[0] = int(rng.integers(100, 140))   # rows wider than some work-groups
    bc = tuple(FACES[int(rng.integers(4))] if rng.random() < 0.35 else orc.FLAG_OBSTACLE for _ in range(6))
    tau = float(rng.uniform(0.5005, 1.2))
    les = rng.random() < 0.5
    forcing = int(rng.integers(3)) - 1      # -1: none, 0: the reference's, 1: body force
    g = tuple(float(x) for x in rng.uniform(-2e-5, 2e-5, 3)) if forcing >= 0 else (0.0, 0.0, 0.0)
    box = None
    if rng.random() < 0.5:
        size = [int(rng.integers(1, max(2, d // 3))) for d in dims]
        origin = [int(rng.integers(1, d - s)) for d, s in zip(dims, size)]
        box = (tuple(origin), tuple(size))
    return dict(dims=tuple(dims), bc=bc, tau=tau, cs=float(rng.uniform(0.1, 0.2)) if les else 0.0,
                lid=float(rng.uniform(0.0, 0.1)), lid_layer=bool(rng.random() < 0.7),
                forcing=max(forcing, 0), g=g, storage=(_capi.STORAGE_F32, _capi.STORAGE_F64)[int(rng.integers(2))],
                wg=(64, 128, 256, 512)[int(rng.integers(4))],
                variant=VARIANTS[int(rng.integers(len(VARIANTS)))] | (0, 0, 4, 32)[int(rng.integers(4))],
                amp=float(rng.choice([0.0, 1e-3, 1e-2])), seed=int(rng.integers(1 << 30)),
                steps=int(rng.integers(1, 10)), box=box)


@pytest.mark.parametrize("case", range(48))
def test_random_domain_equals_oracle(case):
    c = draw(case)
    say(f"case {case}: {c}")
    o, s = make_pair(c["dims"], c["bc"], inv_tau=1.0 / c["tau"], tau=c["tau"], cs=c["cs"], lid=c["lid"], g=c["g"],
                     forcing=c["forcing"], storage=c["storage"], wg=c["wg"], lid_layer=c["lid_layer"], amp=c["amp"],
                     seed=c["seed"], variant=c["variant"])
    if c["box"] is not None:
        origin, size = c["box"]
        o.set_flags(origin, size, orc.FLAG_OBSTACLE)
        s.set(_capi.FIELD_FLAGS, np.full(size[::-1], orc.FLAG_OBSTACLE, np.int32), origin, size)
    o.run(c["steps"])
    s.run(c["steps"])
    np.testing.assert_array_equal(gpu_dd(s), o.dd)
    np.testing.assert_array_equal(s.store(_capi.FIELD_VELOCITY).reshape(3, -1), o.vel)
    np.testing.assert_array_equal(s.store(_capi.FIELD_DENSITY).ravel(), o.dens)
    s.close()


NUMS = [(2, 1, 1), (3, 1, 1), (1, 2, 1), (1, 1, 2), (2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2)]


@pytest.mark.parametrize("case", range(24))
def test_random_decomposition_equals_oracle(case):
    """A seeded random decomposition (subdomain sizes, periodic axes, rings of
    one and two included) with the reference's exchange (all 19 populations
    after alpha, the n.e > 0 ones after beta, x then y then z) and a random
    launch shape, against the oracle's decomposition stepped the same way:
    every slot of every subdomain bit for bit after odd and even counts."""
    from tests.test_gpu_parity import build_gpu_group
    from tests.test_oracle import build_oracle_decomposition

    rng = np.random.default_rng(9_000 + case)
    nums = NUMS[int(rng.integers(len(NUMS)))]
    if rng.random() < 0.25:
        nums = (1,) + nums[1:]   # an x ring of one when periodic
    dims = tuple(nums[a] * int(rng.integers(6, 22)) for a in range(3))
    periodic = tuple(bool(rng.random() < 0.4) for _ in range(3))
    kw = dict(inv_tau=1 / 0.55, tau=0.55, smagorinsky_cs=(0.0, 0.17)[int(rng.integers(2))])
    wg = (64, 128, 256, 512)[int(rng.integers(4))]
    variant = VARIANTS[int(rng.integers(len(VARIANTS)))] | (0, 0, 4, 32)[int(rng.integers(4))]
    alpha_wg, fused = (0, 128, 512)[int(rng.integers(3))], (0, 1)[int(rng.integers(2))]
    say(f"case {case}: dims {dims} nums {nums} periodic {periodic} cs {kw['smagorinsky_cs']} wg {wg} "
        f"variant {variant} alpha {alpha_wg} fused {fused}")
    multi, _ = build_oracle_decomposition(dims, nums, lid=0.06, amp=1e-3, seed=case, periodic=periodic, **kw)
    grp, _, _ = build_gpu_group(dims, nums, lid=0.06, amp=1e-3, seed=case, mode=0, periodic=periodic, **kw)
    grp.set_launch(wg, variant)
    grp.set_alpha_work_group(alpha_wg)
    grp.set_fused_x(fused)
    for steps in (3, 4):
        multi.run(steps)
        grp.step(steps)
        grp.wait()
        for k, (sub, s) in enumerate(zip(multi.subs, grp.solvers)):
            np.testing.assert_array_equal(gpu_dd(s), sub["domain"].dd, err_msg=f"member {k}")
    grp.close()
    for s in grp.solvers:
        s.close()
