"""Serving loop: continuous batching of cavity and channel cases.

``Server`` wraps the native scheduler of the C-ABI (lbm_server_*): cases are
admitted against a device-memory budget into one arena, stepped per storage
type as one fused, graph-replayed launch sequence, frozen at their own step
count and read back while the next chunk runs; requests that arrive while
others are on the device join at the next chunk boundary.

``request_stream`` draws the workload (seeded): Poisson arrivals at a given
rate, cavity or channel with equal odds, cubic grids with edge
round(exp U[ln 32, ln 192]), step counts uniform in 500..5000, storage drawn
from the given set.  ``serve`` replays a stream against a server on the
server's clock and reports cases/s and p50/p99 latency (arrival to result on
the host).

    python -m turbulent_lbm_multigpu_amd.serving --rate 4 --requests 64
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _capi
from ._capi import LbmError

STORAGE_NAMES = {"f32": _capi.STORAGE_F32, "f16": _capi.STORAGE_F16_SHIFTED, "bf16": _capi.STORAGE_BF16_SHIFTED,
                 "f64": _capi.STORAGE_F64}


# Resident-cell limit of the batched legs (lbm_server_config.max_active_cells):
# the queue is admitted in order while the resident cases hold fewer cells.
# 16 Mi cells keep the fused launch filling the device while the largest cases
# (192^3 = 7.1 M cells) run beside few others; measured on the 100-request
# stream against no limit: FP16 at saturation 19.1 vs 18.6 cases/s with the
# mean latency 25 % lower, the mixed stream at 12/s p50 0.34 vs 0.52 s and
# p99 2.23 vs 2.76 s.
BATCH_CELLS = 16 << 20


def make_case(kind: str = "cavity", dims=(32, 32, 32), steps: int = 100, storage: str = "f32", *, tau: float = 0.52,
              cs: float = 0.17, lid: float = 0.05, body_force: float = 1e-6, fill: float = 0.05, r_min: float = 3.0,
              r_max: float = 8.0, perturbation: float = 0.0, seed: int = 1) -> _capi.Case:
    lib = _capi.load_library()
    c = _capi.Case()
    lib.lbm_case_default(C.byref(c))
    c.kind = _capi.CASE_CHANNEL if kind == "channel" else _capi.CASE_CAVITY
    for a in range(3):
        c.dims[a] = int(dims[a])
    c.steps = int(steps)
    c.storage = STORAGE_NAMES[storage] if isinstance(storage, str) else int(storage)
    c.tau = float(tau)
    c.smagorinsky_cs = float(cs)
    c.lid_velocity = float(lid)
    c.body_force = float(body_force)
    c.obstacle_fill = float(fill)
    c.r_min = float(r_min)
    c.r_max = float(r_max)
    c.perturbation = float(perturbation)
    c.seed = int(seed)
    return c


def case_bytes(case: _capi.Case) -> int:
    lib = _capi.load_library()
    n = C.c_size_t()
    rc = lib.lbm_server_case_bytes(C.byref(case), C.byref(n))
    if rc != 0:
        raise LbmError(rc, lib.lbm_last_error().decode())
    return n.value


class Server:
    """The native continuous-batching scheduler (one device)."""

    def __init__(self, memory_budget: float = 8e9, *, max_active: int = 0, chunk_ms: float = 25.0,
                 max_chunk_steps: int = 1000, work_group_size: int = 256, kernel_variant: int = 0,
                 autotune: bool = False, backfill: bool = True, max_active_cells: int = 0, device: int = 0):
        self._lib = _capi.load_library()
        cfg = _capi.ServerConfig()
        self._lib.lbm_server_config_default(C.byref(cfg))
        cfg.device = int(device)
        cfg.memory_budget = int(memory_budget)
        cfg.max_active = int(max_active)
        cfg.chunk_ms = float(chunk_ms)
        cfg.max_chunk_steps = int(max_chunk_steps)
        cfg.work_group_size = int(work_group_size)
        cfg.kernel_variant = int(kernel_variant)
        cfg.autotune = int(bool(autotune))
        cfg.backfill = int(bool(backfill))
        cfg.max_active_cells = int(max_active_cells)
        h = C.c_void_p()
        rc = self._lib.lbm_server_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise LbmError(rc, f"server creation failed (budget {memory_budget:.3g} bytes): "
                               f"{self._lib.lbm_last_error().decode()}")
        self._h = h
        self.cells: Dict[int, tuple] = {}

    def _check(self, rc):
        if rc != 0:
            raise LbmError(rc, self._lib.lbm_server_error(self._h).decode())

    def time(self) -> float:
        return self._lib.lbm_server_time(self._h)

    def submit(self, case: _capi.Case) -> int:
        i = C.c_longlong()
        self._check(self._lib.lbm_server_submit(self._h, C.byref(case), C.byref(i)))
        self.cells[i.value] = (case.dims[2], case.dims[1], case.dims[0])
        return i.value

    def progress(self) -> int:
        n = C.c_int()
        self._check(self._lib.lbm_server_progress(self._h, C.byref(n)))
        return n.value

    def next_done(self) -> Optional[dict]:
        st = _capi.CaseStatus()
        rc = self._lib.lbm_server_next_done(self._h, C.byref(st))
        if rc == _capi.LBM_ERR_STATE:
            return None
        self._check(rc)
        return {k: getattr(st, k) for k, _ in st._fields_}

    def status(self, case_id: int) -> dict:
        st = _capi.CaseStatus()
        self._check(self._lib.lbm_server_status(self._h, int(case_id), C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def fetch(self, case_id: int, keep: bool = True):
        """rho (Z, Y, X) and j (3, Z, Y, X) float32 of a completed case; the
        server releases them (keep=False only releases)."""
        shape = self.cells.pop(case_id)
        if not keep:
            self._check(self._lib.lbm_server_fetch(self._h, int(case_id), None, None))
            return None
        rho = np.empty(shape, np.float32)
        j = np.empty((3,) + shape, np.float32)
        self._check(self._lib.lbm_server_fetch(self._h, int(case_id), rho.ctypes.data_as(C.c_void_p),
                                               j.ctypes.data_as(C.c_void_p)))
        return rho, j

    def stats(self) -> dict:
        st = _capi.ServerStats()
        self._check(self._lib.lbm_server_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def busy(self) -> bool:
        st = self.stats()
        return st["queued"] + st["active"] + st["ready"] > 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lbm_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Request:
    arrival: float
    kind: str
    edge: int
    steps: int
    storage: str
    seed: int
    lid: float = 0.05
    tau: float = 0.52

    def case(self, storage: Optional[str] = None) -> _capi.Case:
        return make_case(self.kind, (self.edge,) * 3, self.steps, storage or self.storage, tau=self.tau,
                         lid=self.lid, seed=self.seed)


def request_stream(n: int, rate: float, seed: int = 20260101, storages: Sequence[str] = ("f32", "f16"),
                   edge_range=(32, 192), steps_range=(500, 5000)) -> List[Request]:
    """Seeded request stream; rate <= 0 or inf puts every arrival at t = 0."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    lo, hi = math.log(edge_range[0]), math.log(edge_range[1])
    for k in range(n):
        if rate and math.isfinite(rate) and rate > 0:
            t += float(rng.exponential(1.0 / rate))
        kind = "channel" if rng.random() < 0.5 else "cavity"
        edge = int(round(math.exp(rng.uniform(lo, hi))))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        storage = str(storages[int(rng.integers(len(storages)))])
        lid = float(rng.uniform(0.02, 0.1))
        out.append(Request(t, kind, edge, steps, storage, seed * 1000003 + k, lid=lid))
    return out


@dataclass
class ServeResult:
    latencies: List[float] = field(default_factory=list)
    wall_s: float = 0.0
    completed: int = 0
    failed: int = 0
    errors: List[str] = field(default_factory=list)
    stats: dict = field(default_factory=dict)
    results: Dict[int, tuple] = field(default_factory=dict)

    def summary(self) -> dict:
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"completed": self.completed, "failed": self.failed, "wall_s": round(self.wall_s, 3),
                "cases_per_s": round(self.completed / self.wall_s, 3) if self.wall_s > 0 else 0.0,
                "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
                "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
                "mean_latency_s": round(float(lat.mean()), 4),
                "mlups": round(self.stats.get("steps_cells", 0) / self.wall_s / 1e6, 1) if self.wall_s > 0 else 0.0,
                "device_busy": round(self.stats.get("device_ms", 0.0) / 1e3 / self.wall_s, 3)
                if self.wall_s > 0 else 0.0,
                "peak_bytes": int(self.stats.get("bytes_peak", 0)), "chunks": int(self.stats.get("chunks", 0))}


def serve(server: Server, requests: Sequence[Request], *, storage: Optional[str] = None, keep_results: bool = False,
          log=None) -> ServeResult:
    """Replay `requests` (arrival offsets in seconds) on the server's clock.
    Latency = time the result reached the host - scheduled arrival."""
    pending = deque(sorted(requests, key=lambda r: r.arrival))
    ids: Dict[int, int] = {}
    arrivals: Dict[int, float] = {}
    res = ServeResult()
    t0 = server.time()
    first_arrival = t0 + (pending[0].arrival if pending else 0.0)
    last_done = first_arrival
    outstanding = 0
    k = 0
    while pending or outstanding:
        now = server.time() - t0
        while pending and pending[0].arrival <= now:
            r = pending.popleft()
            cid = server.submit(r.case(storage))
            ids[cid] = k
            arrivals[cid] = t0 + r.arrival
            k += 1
            outstanding += 1
        if server.busy():
            server.progress()
        elif pending:
            time.sleep(max(0.0, min(pending[0].arrival - (server.time() - t0), 0.01)))
        while True:
            st = server.next_done()
            if st is None:
                break
            cid = st["id"]
            outstanding -= 1
            if st["state"] == _capi.CASE_FAILED:
                res.failed += 1
                try:
                    server.fetch(cid, keep=False)
                except LbmError as e:
                    res.errors.append(str(e))
                continue
            res.latencies.append(st["done_s"] - arrivals[cid])
            last_done = max(last_done, st["done_s"])
            out = server.fetch(cid, keep=keep_results)
            if keep_results:
                res.results[ids[cid]] = out
            res.completed += 1
            if log:
                log(f"  case {ids[cid]} done: {st['cells']} cells x {st['steps_done']} steps, "
                    f"latency {res.latencies[-1]:.3f} s")
        if outstanding and not pending and not server.busy():
            raise RuntimeError(f"server idle with {outstanding} cases outstanding")
    res.wall_s = last_done - first_arrival
    res.stats = server.stats()
    return res


def compare(requests: Sequence[Request], budget: float, *, chunk_ms: float = 25.0, work_group_size: int = 256,
            kernel_variant: int = 0, autotune: bool = False, max_active_cells: int = BATCH_CELLS, log=None) -> dict:
    """The plain FP32 path (one case at a time, FP32) against FP16-shifted
    storage with continuous batching, on the same requests."""
    out = {}
    for name, storage, max_active in (("fp32_one_at_a_time", "f32", 1), ("fp16_batched", "f16", 0)):
        with Server(budget, max_active=max_active, chunk_ms=chunk_ms, work_group_size=work_group_size,
                    kernel_variant=kernel_variant, autotune=autotune,
                    max_active_cells=max_active_cells if max_active == 0 else 0) as srv:
            r = serve(srv, requests, storage=storage, log=log)
        out[name] = r.summary()
        if log:
            log(f"{name}: {json.dumps(out[name])}")
    a, b = out["fp32_one_at_a_time"], out["fp16_batched"]
    out["speedup_cases_per_s"] = round(b["cases_per_s"] / a["cases_per_s"], 3) if a["cases_per_s"] else None
    return out


# The legs of the serving evaluation: (name, storage override or None for
# each request's own, max_active: 1 = one case at a time, 0 = continuous batching)
LEGS = {
    "fp32_one_at_a_time": ("f32", 1),    # the plain FP32 path
    "fp16_one_at_a_time": ("f16", 1),    # shifted storage alone
    "fp32_batched": ("f32", 0),          # continuous batching alone
    "fp16_batched": ("f16", 0),          # both
    "mixed_batched": (None, 0),          # each request's own storage (FP32 or FP16), batched
}


def evaluate(requests: Sequence[Request], budget: float, legs: Sequence[str] = tuple(LEGS), *,
             chunk_ms: float = 25.0, work_group_size: int = 256, kernel_variant: int = 0, autotune: bool = False,
             max_active_cells: int = BATCH_CELLS, log=None) -> dict:
    """Serve the same requests once per leg (see LEGS) and report each leg's
    summary plus its cases/s relative to the plain FP32 path.  The batched
    legs admit at most `max_active_cells` resident cells (0: no limit)."""
    out = {}
    for name in legs:
        storage, max_active = LEGS[name]
        with Server(budget, max_active=max_active, chunk_ms=chunk_ms, work_group_size=work_group_size,
                    kernel_variant=kernel_variant, autotune=autotune,
                    max_active_cells=max_active_cells if max_active == 0 else 0) as srv:
            r = serve(srv, requests, storage=storage)
        out[name] = r.summary()
        if log:
            log(f"  {name}: {json.dumps(out[name])}")
    base = out.get("fp32_one_at_a_time", {}).get("cases_per_s")
    if base:
        out["cases_per_s_vs_fp32_one_at_a_time"] = {
            k: round(v["cases_per_s"] / base, 3) for k, v in out.items() if isinstance(v, dict) and "cases_per_s" in v}
    return out


def served_error_vs_fp64(requests: Sequence[Request], budget: float, *, storages=("f16", "f32"), sample: int = 4,
                         max_edge: int = 96, log=None) -> dict:
    """Re-serve a sample of the requests (the first `sample` whose edge is at
    most `max_edge`) in FP64 and in each of `storages`, and report the error
    of the served fields against FP64 over the interior cells (boundary layers
    excluded): largest |rho - rho64|, largest and rms |u - u64| / U with U the
    largest |u64|.  The worst case of the sample is reported per storage."""
    picked = [r for r in requests if r.edge <= max_edge][:sample]
    fields = {}
    for storage in ("f64",) + tuple(storages):
        with Server(budget, max_active=0) as srv:
            fields[storage] = serve(srv, picked, storage=storage, keep_results=True).results
    out = {"sample": [{"kind": r.kind, "edge": r.edge, "steps": r.steps} for r in picked]}
    for storage in storages:
        worst = {"rho_max_abs_err": 0.0, "u_max_abs_err_over_U": 0.0, "u_rms_err_over_U": 0.0}
        for k in range(len(picked)):
            r64, j64 = (np.asarray(a, np.float64)[..., 1:-1, 1:-1, 1:-1] for a in fields["f64"][k])
            r, j = (np.asarray(a, np.float64)[..., 1:-1, 1:-1, 1:-1] for a in fields[storage][k])
            u64, u = j64 / r64, j / r
            U = float(np.abs(u64).max())
            worst["rho_max_abs_err"] = max(worst["rho_max_abs_err"], float(np.abs(r - r64).max()))
            worst["u_max_abs_err_over_U"] = max(worst["u_max_abs_err_over_U"], float(np.abs(u - u64).max()) / U)
            worst["u_rms_err_over_U"] = max(worst["u_rms_err_over_U"],
                                            float(np.sqrt(np.mean((u - u64) ** 2))) / U)
        out[storage] = {k: float(f"{v:.3g}") for k, v in worst.items()}
        if log:
            log(f"  served {storage} vs FP64: {out[storage]}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rate", type=float, nargs="+", default=[2.0, 8.0, 0.0],
                    help="arrival rates in cases/s (0 = all at once)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--chunk-ms", type=float, default=25.0)
    ap.add_argument("--seed", type=int, default=20260101)
    ap.add_argument("--edge-max", type=int, default=192)
    ap.add_argument("--steps-max", type=int, default=5000)
    ap.add_argument("--mode", choices=["mixed", "compare"], default="compare")
    ap.add_argument("--kernel-variant", type=int, default=0)
    ap.add_argument("--autotune", action="store_true", help="tune each storage pool's launch shape once")
    ap.add_argument("--max-active-cells", type=int, default=BATCH_CELLS,
                    help="resident-cell limit of continuous batching (0: none)")
    a = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    report = {"workload": f"Poisson cavity/channel requests, cubic edge log-uniform 32-{a.edge_max}, steps "
                          f"500-{a.steps_max}, budget {a.budget_gb} GB", "requests": a.requests, "rates": {}}
    for rate in a.rate:
        reqs = request_stream(a.requests, rate, a.seed, edge_range=(32, a.edge_max), steps_range=(500, a.steps_max))
        log(f"rate {rate}: {len(reqs)} requests")
        if a.mode == "compare":
            report["rates"][str(rate)] = compare(reqs, a.budget_gb * 1e9, chunk_ms=a.chunk_ms,
                                                 kernel_variant=a.kernel_variant, autotune=a.autotune,
                                                 max_active_cells=a.max_active_cells, log=log)
        else:
            with Server(a.budget_gb * 1e9, chunk_ms=a.chunk_ms, kernel_variant=a.kernel_variant,
                        autotune=a.autotune, max_active_cells=a.max_active_cells) as srv:
                report["rates"][str(rate)] = serve(srv, reqs).summary()
    print(json.dumps(report))


if __name__ == "__main__":
    main()
