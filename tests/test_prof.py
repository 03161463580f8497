"""Profile summaries (iss_amd/prof.py) on small hand-made rocprofv3 outputs:
the per-step window of a kernel trace and the traffic calibration."""
import csv
import io
import os
import re
from contextlib import redirect_stdout

import pytest

from iss_amd import prof

TRACE_FIELDS = ["Kind", "Agent_Id", "Queue_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"]


def _trace(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=TRACE_FIELDS)
        w.writeheader()
        for name, t0, t1 in rows:
            w.writerow({"Kind": "KERNEL_DISPATCH", "Agent_Id": "Agent 1", "Queue_Id": 1,
                        "Kernel_Name": name, "Start_Timestamp": t0, "End_Timestamp": t1})


def _table(text):
    """kernel -> (calls per step, ms per step) from the window's table"""
    out = {}
    for line in text.splitlines():
        m = re.match(r"\| `(.+)` \| ([0-9.e+-]+) \| ([0-9.]+) \|", line)
        if m:
            out[m.group(1)] = (float(m.group(2)), float(m.group(3)))
    return out


def _run(*args):
    buf = io.StringIO()
    with redirect_stdout(buf):
        prof.window(*args)
    return buf.getvalue()


def _steps(t, n, extra=()):
    """n steps of yields (1 ms) + trials (2 ms) from time t (ns), then the
    extra dispatches; returns the rows and the end time"""
    rows = []
    for _ in range(n):
        rows.append(("void issdev::k_yield_rows<float, 11, false>(args)", t, t + 1_000_000))
        rows.append(("issdev::k_trials(args)", t + 1_000_000, t + 3_000_000))
        t += 4_000_000
    for name, d in extra:
        rows.append((name, t, t + d))
        t += d + 1000
    return rows, t


def test_window_counts_the_last_steps(tmp_path):
    warm, t = _steps(0, 2)
    timed, t = _steps(t, 3)
    rows = warm + timed
    f = os.path.join(tmp_path, "trace.csv")
    _trace(f, rows[::-1])  # (dispatch order in the file does not matter)
    tab = _table(_run(f, "k_yield_rows<float, 11, false>", "3"))
    assert tab["k_trials"] == (1.0, 2.0)
    assert tab["k_yield_rows<float, 11, false>"] == (1.0, 1.0)


def test_window_stops_before_the_work_after_the_steps(tmp_path):
    steps, t = _steps(0, 4)
    # a feed-down after the steps: decays, then a new pass with its own yields
    decays, t = _steps(t, 0, extra=[("issdev::k_decay_pick(args)", 500_000), ("issdev::k_decay_emit(args)", 700_000)])
    again, _ = _steps(t, 1)
    f = os.path.join(tmp_path, "trace.csv")
    _trace(f, steps + decays + again)
    tab = _table(_run(f, "k_yield_rows<float, 11, false>", "2", "k_decay_pick"))
    assert set(tab) == {"k_trials", "k_yield_rows<float, 11, false>"}
    assert tab["k_trials"] == (1.0, 2.0)
    # without the stop the window runs to the end of the trace
    tab = _table(_run(f, "k_yield_rows<float, 11, false>", "2"))
    assert "k_decay_emit" in tab


def test_window_needs_as_many_anchors_as_steps(tmp_path):
    rows, _ = _steps(0, 2)
    f = os.path.join(tmp_path, "trace.csv")
    _trace(f, rows)
    with pytest.raises(SystemExit):
        _run(f, "k_yield_rows<float, 11, false>", "3")


def test_traffic_scales_by_the_copy(tmp_path):
    """FETCH_SIZE and WRITE_SIZE are scaled so that the largest device copy
    reads and writes the calibration size."""
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    for sub, ctr, copy_kib, k_kib in (("fetch", "FETCH_SIZE", 2 * 262144.0, 1000.0),
                                      ("write", "WRITE_SIZE", 262144.0, 500.0)):
        d = os.path.join(tmp_path, sub)
        os.makedirs(d)
        with open(os.path.join(d, "p_counter_collection.csv"), "w", newline="") as fh:
            w = csv.DictWriter(fh, fieldnames=fields)
            w.writeheader()
            w.writerow({"Kernel_Name": "__amd_rocclr_copyBuffer", "Counter_Name": ctr, "Counter_Value": copy_kib})
            w.writerow({"Kernel_Name": "issdev::k_trials(args)", "Counter_Name": ctr, "Counter_Value": k_kib})
    out = os.path.join(tmp_path, "traffic.json")
    with redirect_stdout(io.StringIO()):
        prof.traffic(os.path.join(tmp_path, "fetch"), os.path.join(tmp_path, "write"), out)
    import json
    t = json.load(open(out))["k_trials"]
    assert t["fetch_bytes"] == pytest.approx(1000.0 * 1024 / 2)
    assert t["write_bytes"] == pytest.approx(500.0 * 1024)
