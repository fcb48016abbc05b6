"""CPU checks of the profiler output (the reference's event-file schema) and
of the counter-to-HBM-traffic reduction."""
import csv
import os

import pytest

from turbulent_lbm_multigpu_amd import profiling

TRACE_FIELDS = ["Kernel_Name", "Start_Timestamp", "End_Timestamp"]
PMC_FIELDS = ["Kernel_Name", "Counter_Name", "Counter_Value"]


def write_csv(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_event_ini_roundtrip_and_overlap(tmp_path):
    trace = tmp_path / "kernel_trace.csv"
    write_csv(trace, TRACE_FIELDS, [
        {"Kernel_Name": "step_kernel<0, true, 0>", "Start_Timestamp": 1000, "End_Timestamp": 3000},
        {"Kernel_Name": "exchange_kernel<4u>", "Start_Timestamp": 2500, "End_Timestamp": 2600},
        {"Kernel_Name": "step_kernel<0, false, 0>", "Start_Timestamp": 4000, "End_Timestamp": 5500},
    ])
    ev = profiling.events_from_kernel_trace(str(trace))
    assert [e.uuid for e in ev] == [1, 2, 3]
    assert ev[0].duration == pytest.approx(0.002)   # ms, as CProfilerEvent stores it
    ini = tmp_path / "profile_1_0.ini"
    profiling.write_event_ini(ev, str(ini), total_num_proc=8, current_proc_id=3)
    text = ini.read_text()
    assert text.startswith("[METADATA]\nTOTAL_NUM_PROC : 8\nCURRENT_PROC_ID : 3\n")
    assert "[EVENT2]\nTYPE : DEVICE_KERNEL\nNAME : exchange_kernel<4u>\n" in text
    back = profiling.read_event_ini(str(ini))
    assert [(e.name, e.start, e.end) for e in back] == [(e.name, e.start, e.end) for e in ev]
    pairs = profiling.overlapping(back)
    assert [(a.uuid, b.uuid) for a, b in pairs] == [(1, 2)]
    only_step = profiling.events_from_kernel_trace(str(trace), "step_kernel")
    assert len(only_step) == 2


def test_pmc_traffic_scaling(tmp_path):
    gib = float(1 << 30)
    fetch = tmp_path / "fetch.csv"
    write_csv(fetch, PMC_FIELDS, [
        {"Kernel_Name": "stream_copy4_kernel<true>(x)", "Counter_Name": "FETCH_SIZE", "Counter_Value": gib / 1024 / 2},
        {"Kernel_Name": "lbm::step_kernel<0, true, 0>(y)", "Counter_Name": "FETCH_SIZE", "Counter_Value": 1000.0},
        {"Kernel_Name": "lbm::step_kernel<0, true, 0>(y)", "Counter_Name": "FETCH_SIZE", "Counter_Value": 1200.0},
    ])
    write = tmp_path / "write.csv"
    write_csv(write, PMC_FIELDS, [
        {"Kernel_Name": "stream_copy4_kernel<true>(x)", "Counter_Name": "WRITE_SIZE", "Counter_Value": gib / 1024},
        {"Kernel_Name": "lbm::step_kernel<0, true, 0>(y)", "Counter_Name": "WRITE_SIZE", "Counter_Value": 500.0},
    ])
    t = profiling.pmc_traffic(str(fetch), str(write), "step_kernel", "stream_copy4_kernel<true>", gib)
    assert t["fetch_scale"] == pytest.approx(2.0)
    assert t["write_scale"] == pytest.approx(1.0)
    k = t["kernels"]["lbm::step_kernel<0, true, 0>(y)"]
    assert k["read_bytes"] == pytest.approx(1100.0 * 1024 * 2)
    assert k["write_bytes"] == pytest.approx(500.0 * 1024)
    assert t["mean_bytes_per_launch"] == pytest.approx(k["bytes"])
