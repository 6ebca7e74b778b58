"""The summaries kept under profiles/ are written by periodic-lbm_amd/rocprof_summary.py
from rocprofv3 CSV files; these tests run each mode on small hand-made CSV files whose
figures are known."""
from __future__ import annotations

import csv
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
TOOL = ROOT / "periodic-lbm_amd" / "rocprof_summary.py"


def run(*args):
    subprocess.run([sys.executable, str(TOOL), *map(str, args)], check=True, timeout=60)


def write_csv(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=list(rows[0]))
        w.writeheader()
        w.writerows(rows)


def trace_rows(spec):
    """spec: (name, start_ns, duration_ns, grid) -> kernel_trace.csv rows."""
    return [{"Kernel_Name": n, "Start_Timestamp": t, "End_Timestamp": t + d, "Grid_Size_X": g}
            for n, t, d, g in spec]


def test_trace_mode_averages_the_last_dispatches_of_one_grid(tmp_path):
    spec = [("k_multi<1, 0, 8>", 1000 * i, 500, 700416) for i in range(5)]
    spec += [("k_multi<1, 0, 8>", 10 ** 6, 9000, 352256)]  # another grid: excluded
    write_csv(tmp_path / "t.csv", trace_rows(spec))
    run("trace", tmp_path / "t.csv", tmp_path / "t.md", "--kernel", "k_multi<1, 0, 8>=1000",
        "--last", 3, "--grid-size", 700416, "--cmd", "c")
    row = [l for l in (tmp_path / "t.md").read_text().splitlines() if "k_multi" in l][0]
    cells = [c.strip() for c in row.split("|")]
    assert cells[2:5] == ["3", "0.5", "0.5"]  # dispatches, avg us, min us
    assert cells[-2] == "2.0"  # 1000 B / 500 ns = 2.0 GB/s


def test_trace_mode_first_dispatches(tmp_path):
    spec = [("k_multi<1, 0, 8>", 1000 * i, 400 + 100 * i, 7) for i in range(6)]
    write_csv(tmp_path / "t.csv", trace_rows(spec))
    run("trace", tmp_path / "t.csv", tmp_path / "t.md", "--kernel", "k_multi<1, 0, 8>=1000",
        "--first", 2, "--cmd", "c")
    text = (tmp_path / "t.md").read_text()
    assert "First 2 dispatches" in text
    row = [l for l in text.splitlines() if "k_multi" in l][0]
    cells = [c.strip() for c in row.split("|")]
    assert cells[2:6] == ["2", "0.5", "0.4", "0.5"]  # dispatches, avg, min, max us


def test_window_mode_sums_the_timed_dispatches(tmp_path):
    spec = [("other", 0, 10, 1), ("k_multi<1, 0, 5>", 100, 4000, 1),
            ("k_multi<1, 0, 6>", 5000, 2000, 1), ("k_multi<1, 0, 6>", 8000, 2200, 1),
            ("k_multi<1, 0, 8>", 11000, 3000, 1)]
    write_csv(tmp_path / "t.csv", trace_rows(spec))
    run("window", tmp_path / "t.csv", tmp_path / "w.md", "--sequence", "k_multi<1, 0, 5>",
        "k_multi<1, 0, 6>", "k_multi<1, 0, 6>", "k_multi<1, 0, 8>", "--skip", 1, "--cmd", "c")
    assert "Timed dispatches: 3, 7.2 us in total." in (tmp_path / "w.md").read_text()


def test_span_mode_cuts_at_idle_gaps(tmp_path):
    # autotune dispatch, 800 us idle, a 2-dispatch run, 100 us idle, the timed 3 dispatches
    spec = [("k_batch_multi<0, 2, 8>", 0, 1000, 1), ("k_batch_multi<0, 2, 6>", 801000, 1000, 1),
            ("k_batch_multi<0, 2, 6>", 802000, 1000, 1)]
    t = 903000
    for _ in range(3):
        spec.append(("k_batch_multi<0, 2, 6>", t, 1000, 1))
        t += 1005  # 5 ns gaps
    write_csv(tmp_path / "t.csv", trace_rows(spec))
    run("span", tmp_path / "t.csv", tmp_path / "s.md", "--kernel", "k_batch_multi<0, 2,",
        "--cells", 1000, "--steps", 10, "--cmd", "c")
    text = (tmp_path / "s.md").read_text()
    assert "3 dispatches" in text
    assert "Span of these dispatches: 0.00 ms" in text  # 3010 ns
    assert "(gaps 0.00 ms)" in text
    assert "1000 cells x 10 steps / span = 3322 MLUPS" in text


def test_traffic_mode_calibrates_against_the_copy_probe(tmp_path):
    def rows(counter, copy_kib, kern_kib):
        return [{"Kernel_Name": "plbm::k_copy", "Counter_Name": counter, "Counter_Value": copy_kib,
                 "Grid_Size": 1},
                {"Kernel_Name": "plbm::k_multi<1, 0, 8>", "Counter_Name": counter,
                 "Counter_Value": kern_kib, "Grid_Size": 700416}]
    # the copy probe moves 4 MiB each way; FETCH_SIZE reads half of it, WRITE_SIZE all
    write_csv(tmp_path / "f.csv", rows("FETCH_SIZE", 2048, 1000))
    write_csv(tmp_path / "w.csv", rows("WRITE_SIZE", 4096, 3000))
    run("traffic", tmp_path / "f.csv", tmp_path / "w.csv", tmp_path / "r.json", "--kernel",
        "k_multi<1, 0, 8>", "--cells", 100, "--bytes-per-lu", 144, "--copy-bytes", 4 << 20,
        "--grid-size", 700416, "--cmd", "c")
    rec = json.loads((tmp_path / "r.json").read_text())
    assert rec["calibration"]["fetch_scale"] == 2.0 and rec["calibration"]["write_scale"] == 1.0
    assert rec["read_bytes_per_launch"] == 2000 * 1024
    assert rec["write_bytes_per_launch"] == 3000 * 1024
    assert rec["algorithmic_bytes_per_launch"] == 14400


def test_sq_mode_reports_valu_busy(tmp_path):
    counters = [{"Kernel_Name": "k_multi<1, 0, 8>", "Counter_Name": c, "Counter_Value": v,
                 "Grid_Size": 7}
                for c, v in (("SQ_INSTS_VALU", 1024 * 2.4e9 * 1e-3 / 4), ("SQ_INSTS_SALU", 1e6),
                             ("SQ_INSTS_LDS", 0), ("SQ_INSTS_VMEM_RD", 0),
                             ("SQ_INSTS_VMEM_WR", 0), ("SQ_WAVES", 64))]
    write_csv(tmp_path / "c.csv", counters)
    write_csv(tmp_path / "t.csv", trace_rows([("k_multi<1, 0, 8>", 0, 2000000, 7)]))
    run("sq", tmp_path / "c.csv", tmp_path / "t.csv", tmp_path / "sq.md", "--kernel",
        "k_multi<1, 0, 8>", "--cmd", "c", "--trace-cmd", "t")
    row = [l for l in (tmp_path / "sq.md").read_text().splitlines() if "k_multi" in l][0]
    assert row.rstrip(" |").endswith("0.50")  # 1 ms of VALU issue in a 2 ms launch


def test_sq_mode_first_dispatches(tmp_path):
    # dispatch 1: 1 ms of VALU issue; dispatch 2 (a later, different pass): 3 ms
    unit = 1024 * 2.4e9 * 1e-3 / 4
    counters = [{"Kernel_Name": "k_multi<1, 0, 8>", "Counter_Name": c, "Counter_Value": v * k,
                 "Grid_Size": 7, "Dispatch_Id": k}
                for k in (1, 3)
                for c, v in (("SQ_INSTS_VALU", unit), ("SQ_INSTS_SALU", 1e6), ("SQ_WAVES", 64))]
    write_csv(tmp_path / "c.csv", counters)
    write_csv(tmp_path / "t.csv", trace_rows([("k_multi<1, 0, 8>", 0, 2000000, 7),
                                              ("k_multi<1, 0, 8>", 5000000, 4000000, 7)]))
    run("sq", tmp_path / "c.csv", tmp_path / "t.csv", tmp_path / "sq.md", "--kernel",
        "k_multi<1, 0, 8>@1", "--cmd", "c", "--trace-cmd", "t")
    row = [l for l in (tmp_path / "sq.md").read_text().splitlines() if "k_multi" in l][0]
    cells = [c.strip() for c in row.split("|")]
    assert cells[2] == f"{unit / 1e6:.2f}" and cells[-3] == "2.0000"  # VALU, launch ms
    assert row.rstrip(" |").endswith("0.50")


def test_traffic_mode_first_dispatches(tmp_path):
    def rows(counter, copy_kib, kern_kibs):
        out = [{"Kernel_Name": "plbm::k_copy", "Counter_Name": counter, "Counter_Value": copy_kib,
                "Grid_Size": 1, "Dispatch_Id": 1}]
        out += [{"Kernel_Name": "plbm::k_multi<1, 0, 8>", "Counter_Name": counter,
                 "Counter_Value": v, "Grid_Size": 700416, "Dispatch_Id": 2 + i}
                for i, v in enumerate(kern_kibs)]
        return out
    # two plain passes first, then three that also store the lagged state (twice the writes)
    write_csv(tmp_path / "f.csv", rows("FETCH_SIZE", 2048, [1000, 1000, 1000, 1000, 1000]))
    write_csv(tmp_path / "w.csv", rows("WRITE_SIZE", 4096, [3000, 3000, 6000, 6000, 6000]))
    run("traffic", tmp_path / "f.csv", tmp_path / "w.csv", tmp_path / "r.json", "--kernel",
        "k_multi<1, 0, 8>", "--cells", 100, "--bytes-per-lu", 144, "--copy-bytes", 4 << 20,
        "--grid-size", 700416, "--first", 2, "--cmd", "c")
    rec = json.loads((tmp_path / "r.json").read_text())
    assert rec["dispatches"] == 2 and rec["write_bytes_per_launch"] == 3000 * 1024
