#!/usr/bin/env python3
"""Benchmark: D3Q19 Smagorinsky LES lid-driven cavity, 256^3 FP32, as 2x2x2
subdomains of 128^3 on one GPU with on-device ghost-layer exchange.

One step = one half of the AA pattern (alpha or beta kernel over all eight
subdomains) plus the ghost exchange that follows it, as in
CController::computeNextStep (src/CController.hpp:385-391).  MLUPS counts
the configured 256^3 cells per step (src/CController.hpp:446-452).

The timed steps are one graph launch, captured before the timed region.
Prints one JSON line (rank 0).  The roofline entry divides the algorithmic
152 B per cell update (19 loads + 19 stores of FP32) by the measured average
duration of the step kernel (events on the group's stream): `frac`; and the
whole step (kernel, exchange, launches) by the same bytes: `frac_end_to_end`.
Its traffic is the HBM bytes per launch from the rocprofv3
FETCH_SIZE/WRITE_SIZE summary in profiles/ when that was taken with the same
launch configuration.  The CPU baseline times the C oracle built for the host
(-O3 -march=native, OpenMP) on a bounded sample of the same workload.  Three
more workloads are reported beside the headline number:
* batching: 512 uneven LES cavities, 1000 steps, one fused launch per
  half-step and one graph launch for all steps, autotuned launch shape,
  against one case at a time;
* precision: the 384x192x192 LES channel stored as FP64, FP32, FP16(f - w)
  and BF16(f - w), 10,000 steps, errors against the FP64 run;
* serving: 100 seeded cavity/channel requests through the continuous-batching
  server, five legs at saturation (storage and batching apart) and three
  Poisson rates with p50/p99 latency, and the served error against FP64.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & % HBM roofline (D3Q19 LES 256³ FP32); batched cases/s + p99 latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 152.0

WORKLOAD = dict(domain=(256, 256, 256), nums=(2, 2, 2), cs=0.17, lid=0.1, tau=0.5005, amp=1e-4,
                seed=20260101)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "step_kernel_256cube_les_2x2x2_hbm_traffic.txt")


def measured_traffic(exchange_mode, kernel_variant, wg, fused_x, wg_alpha=0):
    """HBM bytes per step-kernel launch from the committed counter summary,
    if it was measured with this launch configuration (else None)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    cfg, val = None, None
    for line in open(TRAFFIC_FILE):
        if line.startswith("config:"):
            cfg = dict(kv.split("=") for kv in line.split(":", 1)[1].split())
        elif line.startswith("traffic_bytes_per_launch:"):
            val = float(line.split(":", 1)[1])
    if cfg is None:
        return None
    cfg.setdefault("alpha_work_group", cfg.get("work_group"))
    if cfg != {"exchange_mode": str(exchange_mode), "kernel_variant": str(kernel_variant), "work_group": str(wg),
               "fused_x": str(int(fused_x)), "alpha_work_group": str(wg_alpha or wg)}:
        return None
    return val


def batching_leg(log):
    from turbulent_lbm_multigpu_amd import batch

    cases = batch.make_cavities(512, 20260101)
    b, _ = batch.run_batched(cases, 1000, autotune_pairs=2)
    s, _ = batch.run_sequential(cases, 1000)
    log(f"[bench] batching: {b['cases_per_s']} cases/s batched ({b['tuning_steps_kept']} steps while tuning, "
        f"{b['graph_launches']} graph launch for the other {b['steps'] - b['tuning_steps_kept']}), "
        f"{s['cases_per_s']} one at a time")
    return {"workload": "512 LES cavities, edges U{16..96}, lid U[0.02,0.1], tau lognormal(0.52,0.01) >= 0.5002, "
                        "Cs 0.17, FP32, 1000 steps, seed 20260101",
            "cases_per_s": b["cases_per_s"], "mlups": b["mlups"], "mlups_after_tuning": b["mlups_after_tuning"],
            "wall_s": b["wall_s"], "tune_s": b["tune_s"], "run_s": b["run_s"],
            "tuning_steps_kept": b["tuning_steps_kept"],
            "graph_launches": b["graph_launches"], "eager_half_steps": b["eager_half_steps"],
            "launch": b["launch"], "one_at_a_time_cases_per_s": s["cases_per_s"],
            "speedup": round(b["cases_per_s"] / s["cases_per_s"], 3)}


def serving_leg(log, requests=100, rates=(6.0, 12.0, 24.0), seed=20260101):
    """Config 5: one seeded stream of cavity/channel requests, each with its
    own storage (FP32 or FP16-shifted).  At saturation every leg serves the
    same requests, so the storage and batching contributions show apart; at
    three Poisson rates the plain FP32 path and the mixed stream with
    batching give cases/s and p50/p99 latency; a sample of the served cases is
    re-served in FP64 for the error of the served fields."""
    from turbulent_lbm_multigpu_amd import serving

    # every pool tunes its launch shape on its first chunk (the same tuner)
    sat = serving.evaluate(serving.request_stream(requests, 0.0, seed), 200e9, autotune=True, log=log)
    speed = sat["cases_per_s_vs_fp32_one_at_a_time"]
    log(f"[bench] serving at saturation, cases/s vs the plain FP32 path: {json.dumps(speed)}")
    at_rate = {}
    for rate in rates:
        r = serving.evaluate(serving.request_stream(requests, rate, seed), 200e9,
                             legs=("fp32_one_at_a_time", "mixed_batched"), autotune=True, log=log)
        at_rate[str(rate)] = {leg: {k: r[leg][k] for k in ("cases_per_s", "p50_latency_s", "p99_latency_s",
                                                          "device_busy")}
                              for leg in ("fp32_one_at_a_time", "mixed_batched")}
        log(f"[bench] serving at {rate}/s: {json.dumps(at_rate[str(rate)])}")
    err = serving.served_error_vs_fp64(serving.request_stream(requests, 0.0, seed), 200e9, log=log)
    mix = serving.request_stream(requests, 0.0, seed)
    return {"workload": f"{requests} cavity/channel requests (seed {seed}), cubic edge log-uniform 32-192, steps "
                        "U{500..5000}, storage FP32 or FP16-shifted per request, 200 GB budget, launch shape "
                        "autotuned per storage pool",
            "batched_max_active_cells": serving.BATCH_CELLS,
            "fp16_share_of_requests": round(sum(r.storage == "f16" for r in mix) / len(mix), 3),
            "cases_per_s": sat["fp16_batched"]["cases_per_s"],
            "fp32_one_at_a_time_cases_per_s": sat["fp32_one_at_a_time"]["cases_per_s"],
            "speedup_cases_per_s": speed["fp16_batched"],
            "saturation": {k: sat[k] for k in serving.LEGS},
            "cases_per_s_vs_fp32_one_at_a_time": speed,
            "poisson": at_rate,
            "served_error_vs_fp64": err}


def precision_leg(log, dims=(384, 192, 192), steps=10000):
    """Config 3: the LES channel stored as FP64, FP32, FP16(f - w) and
    BF16(f - w), errors of rho, u and kinetic energy against the FP64 run."""
    from turbulent_lbm_multigpu_amd import precision

    res = precision.run(dims, steps, ("f64", "f32", "f16", "bf16"))
    log(f"[bench] precision: {json.dumps(res)}")
    keep = ("mlups", "kernel_variant", "rho_max_abs_err", "u_max_abs_err_over_U", "u_rms_err_over_U", "ke_rel_err",
            "mass_rel_err", "obstacle_fill", "U_ref")
    return {"workload": f"LES channel {dims[0]}x{dims[1]}x{dims[2]}, periodic x, body force g_x 1e-6, tau 0.51, "
                        f"Cs 0.17, seeded spheres r 3-8 filling 5%, {steps} steps from rest; errors over fluid cells "
                        "against the FP64 run (FP64 mass and kinetic-energy reductions on the GPU)",
            "storages": {k: {q: v[q] for q in keep if q in v} for k, v in res.items()}}


def build_group(exchange_mode=0, graph=True, wg=256, kernel_variant=0, lds_per_block=0, fused_x=True,
                colocated=False, wg_alpha=0):
    import numpy as np

    from turbulent_lbm_multigpu_amd import _capi
    from turbulent_lbm_multigpu_amd import decomposition as dec
    from turbulent_lbm_multigpu_amd.solver import LbmGroup, LbmSolver, colocate

    w = WORKLOAD
    mono = dec.monolithic_size(w["domain"], w["nums"])
    subs = dec.decompose(w["domain"], w["nums"])
    solvers = []
    for sd in subs:
        bc = [sd.bc[0][0], sd.bc[0][1], sd.bc[1][0], sd.bc[1][1], sd.bc[2][0], sd.bc[2][1]]
        solvers.append(LbmSolver(sd.domain.size, bc, inv_tau=1.0 / w["tau"], tau=w["tau"], smagorinsky_cs=w["cs"],
                                 lid_velocity=w["lid"], work_group_size=wg,
                                 kernel_variant=kernel_variant, lds_per_block=lds_per_block))
    if colocated:
        colocate(solvers)   # the eight subdomains side by side in one allocation
    for sd, s in zip(subs, solvers):
        s.init_perturbed(w["amp"], w["seed"], 0, dec.key_origin(sd), mono)
        if sd.has_lid:
            o, sz = dec.lid_region(sd.domain.size)
            s.set(_capi.FIELD_FLAGS, np.full(sz[::-1], _capi.FLAG_VELOCITY_INJECTION, np.int32), o, sz)
    grp = LbmGroup(solvers)
    for sd in subs:
        for c in sd.comms:
            grp.add_comm(sd.uid, c.dst_id, c.send_origin, c.recv_origin, c.send_size, c.comm_direction)
    grp.set_exchange_mode(exchange_mode)
    grp.set_fused_x(fused_x)
    if wg_alpha:
        grp.set_alpha_work_group(wg_alpha)
    grp.set_graph(graph)
    return grp


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(repeats=5, warmup=20, window=100, cavity_steps=1000):
    """The C oracle built for this host (-O3 -march=native, OpenMP over every
    thread the process may use) on the same AA-pattern algorithm:

    * the 256^3 LES cavity as 2x2x2 subdomains: `warmup` steps, then
      `repeats` windows of `window` steps, median MLUPS (the headline's CPU
      counterpart);
    * the 64^3 BGK cavity (tau 0.6, lid 0.05): `cavity_steps` steps in full,
      median of `repeats` runs.

    The timing build allows FMA contraction; the parity build of the tests
    does not.  Both compile the same source."""
    import statistics
    import tempfile

    import numpy as np

    from oracle import build_oracle
    from oracle import oracle as orc
    from turbulent_lbm_multigpu_amd import decomposition as dec

    lib_path = build_oracle.build_timing(tempfile.mkdtemp(prefix="lbm_oracle_timing_"))
    threads = int(orc.load(lib_path).oracle_num_threads())

    w = WORKLOAD
    mono = dec.monolithic_size(w["domain"], w["nums"])
    subs = []
    for sd in dec.decompose(w["domain"], w["nums"]):
        bc = [sd.bc[0][0], sd.bc[0][1], sd.bc[1][0], sd.bc[1][1], sd.bc[2][0], sd.bc[2][1]]
        d = orc.OracleDomain(sd.domain.size, bc, inv_tau=1.0 / w["tau"], tau=w["tau"], smagorinsky_cs=w["cs"],
                             lid_velocity=w["lid"], store_velocity=False, store_density=False, lib_path=lib_path)
        d.reset(amplitude=w["amp"], seed=w["seed"], key_origin=dec.key_origin(sd), key_dims=mono)
        if sd.has_lid:
            o, sz = dec.lid_region(sd.domain.size)
            d.set_flags(o, sz, orc.FLAG_INJECTION)
        subs.append({"domain": d, "comms": [(c.dst_id, c.send_origin, c.recv_origin, c.send_size,
                                             c.comm_direction) for c in sd.comms]})
    multi = orc.OracleMulti(subs)
    multi.run(warmup)
    cells = float(np.prod(w["domain"]))
    les = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        multi.run(window)
        les.append(cells * window / (time.perf_counter() - t0) / 1e6)
    del multi, subs

    cav = []
    for _ in range(repeats):
        dims = (64, 64, 64)
        d = orc.OracleDomain(dims, (1,) * 6, inv_tau=1.0 / 0.6, tau=0.6, lid_velocity=0.05,
                             store_velocity=False, store_density=False, lib_path=lib_path)
        o, sz = dec.lid_region(dims)
        d.set_flags(o, sz, orc.FLAG_INJECTION)
        t0 = time.perf_counter()
        d.run(cavity_steps)
        cav.append(64 ** 3 * cavity_steps / (time.perf_counter() - t0) / 1e6)
    return {
        "value": round(statistics.median(les), 2),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"C oracle built -O3 -march=native with OpenMP ({threads} threads, {_cpu_model()}, "
                  f"{os.cpu_count()} CPUs visible): 256^3 LES as 2x2x2 subdomains, {warmup} warm-up steps then "
                  f"median of {repeats} windows of {window} steps",
        "cpu_model": _cpu_model(),
        "windows_mlups": [round(v, 2) for v in les],
        "cavity_64_bgk_mlups": round(statistics.median(cav), 2),
        "cavity_64_bgk_sample": f"64^3 BGK cavity, tau 0.6, lid 0.05, {cavity_steps} steps in full, median of "
                                f"{repeats} runs",
    }


def copy_probe(lib, nbytes=1 << 30, reps=20):
    """Best read+write bandwidth of a plain device copy (three copy kernels)."""
    import ctypes

    best = 0.0
    for v in range(3):
        gbs, ms = ctypes.c_double(), ctypes.c_double()
        if lib.lbm_device_copy_bandwidth(0, nbytes, reps, v, ctypes.byref(gbs), ctypes.byref(ms)) == 0:
            print(f"[bench] device copy of {nbytes >> 20} MiB, kernel variant {v}: {gbs.value:.0f} GB/s",
                  file=sys.stderr, flush=True)
            best = max(best, gbs.value)
    return best


def arg_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--timing-steps", type=int, default=200)
    ap.add_argument("--exchange-mode", type=int, default=1)
    ap.add_argument("--wg", type=int, default=256)
    ap.add_argument("--wg-alpha", type=int, default=512,
                    help="alpha's work-group size (0: --wg); at 512 its chain tiles hold two rows per subdomain")
    ap.add_argument("--kernel-variant", type=int, default=35,
                    help="non-temporal loads and stores (3), tiles dealt to the XCDs in chunks of 16 (| 32)")
    ap.add_argument("--lds-per-block", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--graph-pairs", type=int, default=1000,
                    help="most beta+alpha pairs one graph launch replays")
    ap.add_argument("--fused-x", type=int, default=1, choices=[0, 1],
                    help="exchange the x faces inside the step kernel")
    ap.add_argument("--colocate", action="store_true",
                    help="move the eight subdomains into one shared allocation (default: allocations of their own)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=100)
    return ap


def main():
    args = arg_parser().parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every subdomain runs on a single device "
              "(no multi-GPU transport); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from turbulent_lbm_multigpu_amd import _capi

    lib = _capi.load_library()
    print(f"[bench] building 256^3 LES 2x2x2 (exchange_mode={args.exchange_mode}, wg={args.wg})",
          file=sys.stderr, flush=True)
    grp = build_group(args.exchange_mode, not args.no_graph, args.wg, args.kernel_variant,
                      args.lds_per_block, bool(args.fused_x), args.colocate, args.wg_alpha)
    grp.set_graph_pairs(args.graph_pairs)
    # the plain-copy reference of the roofline, measured before the step loop
    # (untimed; it also brings the device out of idle)
    copy_gbs = copy_probe(lib)
    grp.step(args.warmup)
    grp.wait()
    # capture the graphs of the timed steps beforehand (recording runs
    # nothing).  Capturing them before the warm-up instead, so that the warm-up
    # runs right before the timed region, measured 1 % slower.
    grp.prepare_steps(args.steps)
    lib.lbm_device_synchronize(0)
    captured = grp.launch_counts()["graphs_captured"]
    t0 = time.perf_counter()
    grp.step(args.steps)
    t_call = time.perf_counter()
    grp.wait()
    lib.lbm_device_synchronize(0)
    t1 = time.perf_counter()
    el = t1 - t0
    cells = 256 ** 3
    mlups = cells * args.steps / el / 1e6
    print(f"[bench] timed {args.steps} steps in {el:.4f} s: {mlups:.0f} MLUPS (the launch call returned after "
          f"{(t_call - t0) * 1e6:.0f} us; graphs captured in the timed region: "
          f"{grp.launch_counts()['graphs_captured'] - captured})", file=sys.stderr, flush=True)

    # per-launch duration of the step kernel, events on the group's stream
    grp.set_timing(True)
    grp.step(args.timing_steps)
    tm = grp.timing()
    grp.set_timing(False)
    kern_ms = tm["step_kernel_ms"]
    achieved = BYTES_PER_CELL * cells / (kern_ms * 1e-3) / 1e9 if kern_ms > 0 else 0.0
    print(f"[bench] step kernel {kern_ms * 1e3:.1f} us/launch (beta {tm['beta_kernel_ms'] * 1e3:.1f}, "
          f"alpha {tm['alpha_kernel_ms'] * 1e3:.1f}), exchange {tm['exchange_ms'] * 1e3:.1f} us/launch",
          file=sys.stderr, flush=True)
    mass, ke = grp.solvers[0].mass_energy()

    def log(m):
        print(m, file=sys.stderr, flush=True)

    traffic = measured_traffic(args.exchange_mode, args.kernel_variant, args.wg, args.fused_x, args.wg_alpha)
    grp.close()
    for s in grp.solvers:
        s.close()
    batching = None if args.no_batch else batching_leg(log)
    precision_res = None if args.no_precision else precision_leg(log)
    serving_res = None if args.no_serving else serving_leg(log, requests=args.serving_requests)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_repeats)
        print(f"[bench] CPU baseline {cpu['value']} MLUPS on {cpu['cores']} threads", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "D3Q19 Smagorinsky LES lid-driven cavity 256^3 FP32, 2x2x2 subdomains of 128^3 "
                        "with on-device ghost exchange",
            "cs": WORKLOAD["cs"], "lid": WORKLOAD["lid"], "tau0": WORKLOAD["tau"],
            "init": "f_eq(1, du), du ~ U(+-1e-4), seed 20260101",
            "exchange_mode": args.exchange_mode, "work_group": args.wg,
            "alpha_work_group": args.wg_alpha or args.wg, "graph": not args.no_graph,
            "kernel_variant": args.kernel_variant, "fused_x": bool(args.fused_x),
            "graph_pairs": args.graph_pairs, "colocated": args.colocate,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(traffic) if traffic else None,
            # the whole step (kernel, exchange and launch gaps): MLUPS x 152 B against 8 TB/s
            "frac_end_to_end": round(mlups * 1e6 * BYTES_PER_CELL / (HBM_PEAK_GBS * 1e9), 4),
        },
        "cpu_baseline": cpu,
        "step_kernel_us": round(kern_ms * 1e3, 2),
        "beta_kernel_us": round(tm["beta_kernel_ms"] * 1e3, 2),
        "alpha_kernel_us": round(tm["alpha_kernel_ms"] * 1e3, 2),
        "copy_probe_gbs": round(copy_gbs, 1),
        "exchange_kernel_us": round(tm["exchange_ms"] * 1e3, 2),
        "subdomain0_mass_ke": [mass, ke],
        "batching": batching,
        "precision": precision_res,
        "serving": serving_res,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
