"""Profiler output in the reference's event-file schema, and HBM traffic from
rocprofv3 counters.

* ``events_from_kernel_trace`` reads a rocprofv3 ``--kernel-trace
  --output-format csv`` file; ``write_event_ini`` writes the events as the
  reference's profiler does (``[METADATA]`` with TOTAL_NUM_PROC /
  CURRENT_PROC_ID, src/CController.hpp:503-517, then one ``[EVENTn]`` per
  event with TYPE, NAME, START/END in ns and DURATION in ms,
  src/libtools/CProfilerEvent.hpp:81-91, src/libtools/CProfiler.hpp:40-48),
  so the reference's overlap analysis (profile.py:36-58) reads them;
  ``overlapping`` is that analysis.
* ``pmc_traffic`` turns two ``--pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs
  into HBM bytes per launch of a kernel, each counter scaled by what it reads
  for a copy kernel of known size in the same run.

    python -m turbulent_lbm_multigpu_amd.profiling ini <kernel_trace.csv> <out.ini>
    python -m turbulent_lbm_multigpu_amd.profiling traffic <fetch.csv> <write.csv> \\
        --kernel step_kernel --copy 'stream_copy4_kernel<true>' --copy-bytes 1073741824
"""
from __future__ import annotations

import argparse
import csv
import json
import sys
from collections import defaultdict
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence


@dataclass
class ProfilerEvent:
    uuid: int
    event_type: str
    name: str
    start: int      # ns
    end: int        # ns
    duration: float  # ms

    def overlap(self, other: "ProfilerEvent") -> bool:
        return self.start < other.end and self.end > other.start


def events_from_kernel_trace(path: str, name_filter: str = "") -> List[ProfilerEvent]:
    out = []
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            name = row["Kernel_Name"]
            if name_filter and name_filter not in name:
                continue
            s, e = int(row["Start_Timestamp"]), int(row["End_Timestamp"])
            out.append(ProfilerEvent(len(out) + 1, "DEVICE_KERNEL", name, s, e, (e - s) / 1e6))
    return out


def write_event_ini(events: Sequence[ProfilerEvent], path: str, total_num_proc: int = 1, current_proc_id: int = 0):
    with open(path, "w") as f:
        f.write(f"[METADATA]\nTOTAL_NUM_PROC : {total_num_proc}\nCURRENT_PROC_ID : {current_proc_id}\n\n")
        for ev in events:
            f.write(f"[EVENT{ev.uuid}]\nTYPE : {ev.event_type}\nNAME : {ev.name}\n"
                    f"# start/end in nanoseconds\nSTART : {ev.start}\nEND : {ev.end}\n"
                    f"# duration in milliseconds\nDURATION : {ev.duration}\n\n")


def read_event_ini(path: str) -> List[ProfilerEvent]:
    events, cur = [], None
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        if line.startswith("[EVENT"):
            cur = {"uuid": int(line[6:-1])}
            events.append(cur)
        elif line.startswith("["):
            cur = None
        elif cur is not None and " : " in line:
            k, v = line.split(" : ", 1)
            cur[k] = v
    return [ProfilerEvent(e["uuid"], e["TYPE"], e["NAME"], int(e["START"]), int(e["END"]), float(e["DURATION"]))
            for e in events]


def overlapping(events: Sequence[ProfilerEvent]):
    """Pairs of events whose [start, end) intervals intersect (profile.py:36-58)."""
    return [(a, b) for i, a in enumerate(events) for b in events[i + 1:] if a.overlap(b)]


def counter_means(path: str) -> Dict[str, float]:
    """Mean counter value (KiB per dispatch for *_SIZE) per kernel name."""
    acc = defaultdict(list)
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            acc[row["Kernel_Name"]].append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def counter_table(paths: Sequence[str], kernel: str = "") -> Dict[str, Dict[str, float]]:
    """Mean value per dispatch of every counter, per kernel name (kernels whose
    name contains `kernel`), from one or more rocprofv3 counter_collection
    CSV files (one --pmc pass each)."""
    acc = defaultdict(lambda: defaultdict(list))
    for path in paths:
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                if kernel in row["Kernel_Name"]:
                    acc[row["Kernel_Name"]][row["Counter_Name"]].append(float(row["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def pmc_traffic(fetch_csv: str, write_csv: str, kernel: str, copy: str, copy_bytes: float,
                fetch_scale: Optional[float] = None, write_scale: Optional[float] = None) -> dict:
    """HBM bytes per launch of every kernel whose name contains `kernel`.
    The copy kernel (`copy` in its name) reads and writes `copy_bytes` per
    launch; its FETCH_SIZE and WRITE_SIZE readings give the scale factors,
    unless they are given (a run without the copy kernel, scaled by another
    run's calibration on the same part)."""
    fetch, write = counter_means(fetch_csv), counter_means(write_csv)

    def pick(table, key):
        hits = {k: v for k, v in table.items() if key in k}
        if not hits:
            raise KeyError(f"no kernel matching {key!r}")
        return hits

    if fetch_scale is None or write_scale is None:
        cf = list(pick(fetch, copy).values())[0]
        cw = list(pick(write, copy).values())[0]
    f_scale = fetch_scale if fetch_scale is not None else copy_bytes / (cf * 1024.0)
    w_scale = write_scale if write_scale is not None else copy_bytes / (cw * 1024.0)
    out = {"fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for name, kib in pick(fetch, kernel).items():
        wk = write.get(name)
        if wk is None:
            continue
        rb, wb = kib * 1024.0 * f_scale, wk * 1024.0 * w_scale
        out["kernels"][name] = {"fetch_kib": kib, "write_kib": wk, "read_bytes": rb, "write_bytes": wb,
                                "bytes": rb + wb}
    ks = out["kernels"].values()
    out["mean_bytes_per_launch"] = sum(k["bytes"] for k in ks) / max(len(out["kernels"]), 1)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("ini")
    a.add_argument("trace")
    a.add_argument("out")
    a.add_argument("--filter", default="")
    b = sub.add_parser("traffic")
    b.add_argument("fetch")
    b.add_argument("write")
    b.add_argument("--kernel", default="step_kernel")
    b.add_argument("--copy", default="stream_copy4_kernel<true>")
    b.add_argument("--copy-bytes", type=float, default=float(1 << 30))
    b.add_argument("--fetch-scale", type=float, default=None, help="use this FETCH_SIZE scale (no copy kernel)")
    b.add_argument("--write-scale", type=float, default=None, help="use this WRITE_SIZE scale (no copy kernel)")
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--kernel", default="")
    args = ap.parse_args(argv)
    if args.cmd == "counters":
        tab = counter_table(args.csv, args.kernel)
        names = sorted({n for cs in tab.values() for n in cs})
        print("kernel\t" + "\t".join(names))
        for k in sorted(tab):
            print(k + "\t" + "\t".join(f"{tab[k].get(n, float('nan')):.6g}" for n in names))
    elif args.cmd == "ini":
        ev = events_from_kernel_trace(args.trace, args.filter)
        write_event_ini(ev, args.out)
        print(f"{len(ev)} events, {len(overlapping(ev))} overlapping pairs", file=sys.stderr)
    else:
        print(json.dumps(pmc_traffic(args.fetch, args.write, args.kernel, args.copy, args.copy_bytes,
                                     args.fetch_scale, args.write_scale), indent=1))


if __name__ == "__main__":
    main()
