"""Seeded random single domains against the CPU oracle, bit for bit.

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
                variant=VARIANTS[int(rng.integers(len(VARIANTS)))] | (4 if rng.random() < 0.25 else 0),
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
