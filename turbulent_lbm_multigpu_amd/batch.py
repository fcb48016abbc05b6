"""Batching: many independent LES cavities of uneven size in one fused launch.

The cavity set (seeded): every edge uniform in 16..96 cells per axis, lid
speed uniform in [0.02, 0.1], tau = max(0.5002, lognormal(ln 0.52, 0.01)),
Cs = 0.17, FP32, walls on every face and the lid layer at y = ny-2, starting
from f = w_i.  ``run_batched`` puts all cases into one group (one step kernel
launch per half-step over every case's tiles, the beta+alpha pair replayed as
a hipGraph) with the launch shape chosen by the group autotuner;
``run_sequential`` steps each case on its own, one launch per half-step, as a
caller of the plain solver would.

    python -m turbulent_lbm_multigpu_amd.batch --cases 512 --steps 1000
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import _capi
from . import decomposition as dec
from .solver import LbmGroup, LbmSolver, colocate


@dataclass
class Cavity:
    dims: tuple
    lid: float
    tau: float


def make_cavities(n: int = 512, seed: int = 20260101, edge=(16, 96), lid=(0.02, 0.1), tau_median: float = 0.52,
                  tau_sigma: float = 0.01, tau_min: float = 0.5002) -> List[Cavity]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dims = tuple(int(v) for v in rng.integers(edge[0], edge[1] + 1, size=3))
        lv = float(rng.uniform(lid[0], lid[1]))
        tau = max(tau_min, float(rng.lognormal(math.log(tau_median), tau_sigma)))
        out.append(Cavity(dims, lv, tau))
    return out


def build_solver(c: Cavity, *, storage: int = _capi.STORAGE_F32, cs: float = 0.17, work_group_size: int = 256,
                 kernel_variant: int = 0) -> LbmSolver:
    s = LbmSolver(c.dims, (_capi.FLAG_OBSTACLE,) * 6, inv_tau=1.0 / c.tau, tau=c.tau, smagorinsky_cs=cs,
                  lid_velocity=c.lid, storage=storage, work_group_size=work_group_size,
                  kernel_variant=kernel_variant)
    origin, size = dec.lid_region(c.dims)
    s.set(_capi.FIELD_FLAGS, np.full(size[::-1], _capi.FLAG_VELOCITY_INJECTION, np.int32), origin, size)
    return s


def _restart(solvers: Sequence[LbmSolver], cases: Sequence[Cavity]):
    """Every case back to its initial state (f = w_i, walls, lid layer)."""
    for s, c in zip(solvers, cases):
        s.reset()
        origin, size = dec.lid_region(c.dims)
        s.set(_capi.FIELD_FLAGS, np.full(size[::-1], _capi.FLAG_VELOCITY_INJECTION, np.int32), origin, size)


def run_batched(cases: Sequence[Cavity], steps: int, *, autotune_pairs: int = 2, storage: int = _capi.STORAGE_F32,
                keep: bool = False, colocated: bool = False, log=None):
    """All cases in one group (their buffers moved side by side into one
    allocation if `colocated`).  The autotuner times real steps of the
    batch (16 launch shapes, then alpha at 3 work-group sizes of its own,
    then the XCD tile map switched, (pairs + 1) step pairs each): every launch shape computes the same
    arithmetic, so those are `steps` of every case like any other, and the
    rest run as one graph launch (one fused step-kernel node per
    half-step).  If tuning took more steps than `steps`, the cases restart
    from their initial state and all `steps` run in the graph.  Tuning is
    counted in the wall time."""
    solvers = [build_solver(c, storage=storage) for c in cases]
    if colocated:
        colocate(solvers)
    g = LbmGroup(solvers)
    cells = sum(s.cells for s in solvers)
    t0 = time.perf_counter()
    tuned = None
    done = 0   # steps every case has taken
    if autotune_pairs > 0:
        tuned = g.autotune(autotune_pairs)   # whole pairs from a beta step: the group's parity stays even
        g.wait()
        done = solvers[0].step_counter
        if done > steps:
            _restart(solvers, cases)
            done = 0
    rest = steps - done
    g.set_graph_pairs(max(1, (rest + 1) // 2))
    g.prepare_steps(rest)
    t1 = time.perf_counter()
    before = g.launch_counts()
    g.step(rest)
    g.wait()
    t2 = time.perf_counter()
    after = g.launch_counts()
    res = {"cases": len(cases), "steps": steps, "tuning_steps_kept": done, "cells": cells, "colocated": colocated,
           "tune_s": round(t1 - t0, 4),
           "run_s": round(t2 - t1, 4), "wall_s": round(t2 - t0, 4),
           "cases_per_s": round(len(cases) / (t2 - t0), 3),
           "mlups": round(cells * steps / (t2 - t0) / 1e6, 1),
           "mlups_after_tuning": round(cells * rest / (t2 - t1) / 1e6, 1) if rest else None,
           "graph_launches": after["graph_launches"] - before["graph_launches"],
           "eager_half_steps": after["eager_half_steps"] - before["eager_half_steps"],
           "launch": dict(zip(("work_group_size", "kernel_variant"), g.launch()),
                          alpha_work_group=g.alpha_work_group()), "autotune": tuned}
    if log:
        log(f"batched: {json.dumps(res)}")
    fields = [s.moments() for s in solvers] if keep else None
    g.close()
    for s in solvers:
        s.close()
    return res, fields


def run_sequential(cases: Sequence[Cavity], steps: int, *, storage: int = _capi.STORAGE_F32, keep: bool = False,
                   log=None):
    """One case after the other through the plain solver (eager launches)."""
    cells = 0
    fields = [] if keep else None
    total = 0.0
    for c in cases:
        s = build_solver(c, storage=storage)
        t0 = time.perf_counter()
        s.run(steps)
        s.wait()
        total += time.perf_counter() - t0
        cells += s.cells
        if keep:
            fields.append(s.moments())
        s.close()
    res = {"cases": len(cases), "steps": steps, "cells": cells, "wall_s": round(total, 4),
           "cases_per_s": round(len(cases) / total, 3), "mlups": round(cells * steps / total / 1e6, 1)}
    if log:
        log(f"sequential: {json.dumps(res)}")
    return res, fields


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--cases", type=int, default=512)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--autotune-pairs", type=int, default=2)
    ap.add_argument("--sequential", action="store_true", help="also time the one-case-at-a-time path")
    ap.add_argument("--colocate", action="store_true",
                    help="move every case into one shared allocation (default: allocations of their own)")
    a = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    cases = make_cavities(a.cases, a.seed)
    out = {"workload": f"{a.cases} LES cavities, edges U{{16..96}}, lid U[0.02,0.1], tau lognormal(0.52, 0.01) "
                       f">= 0.5002, Cs 0.17, FP32, {a.steps} steps", "batched": run_batched(
                           cases, a.steps, autotune_pairs=a.autotune_pairs, colocated=a.colocate,
                           log=log)[0]}
    if a.sequential:
        out["sequential"] = run_sequential(cases, a.steps, log=log)[0]
        out["speedup"] = round(out["batched"]["cases_per_s"] / out["sequential"]["cases_per_s"], 3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
