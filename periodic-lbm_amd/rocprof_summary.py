"""Summarise rocprofv3 CSV output into the short text records kept under profiles/.

  python periodic-lbm_amd/rocprof_summary.py stats  <kernel_stats.csv> <out.md> --cmd "<cmd>"
  python periodic-lbm_amd/rocprof_summary.py trace  <kernel_trace.csv> <out.md>
         --kernel "<substring>=<bytes per launch>" [--kernel ...] --last N | --first N
         --cmd "<cmd>"
  python periodic-lbm_amd/rocprof_summary.py window <kernel_trace.csv> <out.md>
         --sequence <name substring> ... [--bytes B ...] --skip K --cmd "<cmd>"
  python periodic-lbm_amd/rocprof_summary.py span <kernel_trace.csv> <out.md>
         --kernel <substring> [--kernel ...] [--cells N --steps K] --cmd "<cmd>"
  python periodic-lbm_amd/rocprof_summary.py sq <counter_collection.csv> <kernel_trace.csv>
         <out.md> --kernel <substring>[@first N] [--kernel ...] --grid-size N --cmd "<pmc cmd>"
         --trace-cmd "<trace cmd>"
  python periodic-lbm_amd/rocprof_summary.py traffic <fetch_counter_collection.csv>
         <write_counter_collection.csv> <out.json> --kernel <substring> --cells N
         --bytes-per-lu B [--first N] --cmd "<cmd>"

The pass kernels' first dispatches in bench.py are the autotuner's timing of plain passes
of each depth; later ones of the same kernel can be passes that also store the lagged
state (the autotuner's timing of the last pass of a run, the timed run's last pass).

The traffic record calibrates FETCH_SIZE / WRITE_SIZE (KiB per dispatch) against the
engine's copy probe (k_copy, which reads and writes a known number of bytes in the same
process) and reports HBM bytes per launch of the selected kernel beside its algorithmic
bytes (cells x bytes per lattice update).
"""
from __future__ import annotations

import argparse
import csv
import json
import statistics
from pathlib import Path


def _rows(path):
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


def stats(args):
    rows = _rows(args.csv)
    lines = [f"# rocprofv3 kernel statistics", "", f"Command: `{args.cmd}`", "",
             "| kernel | calls | avg us | min us | max us | share % |",
             "|---|---:|---:|---:|---:|---:|"]
    for r in rows:
        lines.append(f"| `{r['Name'][:90]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                     f"{float(r['Percentage']):.2f} |")
    Path(args.out).write_text("\n".join(lines) + "\n")


def trace(args):
    """Per-kernel durations from kernel_trace.csv, restricted to the last --last dispatches
    of each selected kernel (the timed region of bench.py, after warm-up and tuning), or
    with --first to the first ones (the autotuner's timing of each pass depth, which
    precedes its timing of the passes that also store the lagged state)."""
    rows = _rows(args.csv)
    which = f"First {args.first}" if args.first else f"Last {args.last}"
    lines = [f"# rocprofv3 kernel trace: dispatch durations", "", f"Command: `{args.cmd}`", "",
             f"{which} dispatches of each kernel below"
             + (f" with grid size {', '.join(map(str, args.grid_size))}" if args.grid_size else "")
             + ".", "",
             "| kernel | dispatches | avg us | min us | max us | bytes/launch | GB/s at avg |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for spec in args.kernel:
        sub, nbytes = spec.rsplit("=", 1)
        d = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
              r["Kernel_Name"]) for r in rows if sub in r["Kernel_Name"]
             and (not args.grid_size or int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0)
                  in args.grid_size)]
        d.sort()
        sel = [x[1] for x in (d[:args.first] if args.first else d[-args.last:])]
        if not sel:
            raise SystemExit(f"no dispatch of {sub!r}")
        avg = sum(sel) / len(sel)
        nb = float(nbytes)
        lines.append(f"| `{d[-1][2][:70]}` | {len(sel)} | {avg / 1e3:.1f} | {min(sel) / 1e3:.1f} | "
                     f"{max(sel) / 1e3:.1f} | {nb:.4g} | {nb / avg:.1f} |")
    Path(args.out).write_text("\n".join(lines) + "\n")


def window(args):
    """The dispatches of one stretch of the run, located as the first place in the trace
    where consecutive dispatches match the given kernel-name substrings in order (for
    bench.py's timed region: its warm-up pass, then the timed passes)."""
    rows = _rows(args.csv)
    d = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                r["Kernel_Name"]) for r in rows)
    seq = args.sequence
    start = next((i for i in range(len(d) - len(seq) + 1)
                  if all(seq[j] in d[i + j][2] for j in range(len(seq)))), None)
    if start is None:
        raise SystemExit(f"sequence {seq} not found")
    lines = [f"# rocprofv3 kernel trace: {args.title}", "", f"Command: `{args.cmd}`", "",
             "Consecutive dispatches located by their kernel names (the first "
             f"{args.skip} of them precede the timed region).", "",
             "| # | kernel | us | algorithmic bytes | GB/s |", "|---:|---|---:|---:|---:|"]
    total = 0.0
    for j in range(len(seq)):
        t, dur, name = d[start + j]
        nb = args.bytes[j] if j < len(args.bytes) else 0.0
        timed = j >= args.skip
        total += dur if timed else 0.0
        lines.append(f"| {j}{'' if timed else ' (before)'} | `{name[:60]}` | {dur / 1e3:.1f} | "
                     f"{nb:.4g} | {nb / dur if nb else 0:.1f} |")
    lines += ["", f"Timed dispatches: {len(seq) - args.skip}, {total / 1e3:.1f} us in total."]
    Path(args.out).write_text("\n".join(lines) + "\n")


def _per_kernel(rows, counter, grid_size=None):
    """Counter values per kernel name; with grid_size, only the dispatches of the kernel
    itself with that grid size (the same kernel also runs on other lattices)."""
    agg = {}
    for r in rows:
        if r["Counter_Name"] != counter:
            continue
        agg.setdefault(r["Kernel_Name"], []).append(
            (float(r["Counter_Value"]), int(r.get("Grid_Size") or 0),
             int(r.get("Dispatch_Id") or 0)))
    return agg


def _values(agg, name, grid_size=None, first=0):
    """The kernel's values in dispatch order; with first, only its first dispatches."""
    sel = sorted((d, v) for v, g, d in agg[name] if grid_size is None or g == grid_size)
    return [v for _, v in (sel[:first] if first else sel)]


def _pick(agg, sub, grid_size=None, first=0):
    keys = [k for k in agg if sub in k]
    if not keys:
        raise SystemExit(f"no kernel matching {sub!r} in {list(agg)}")
    vals = _values(agg, keys[0], grid_size, first)
    if not vals:
        raise SystemExit(f"no dispatch of {keys[0]!r} with grid size {grid_size}")
    return keys[0], statistics.median(vals)


def traffic(args):
    fetch = _per_kernel(_rows(args.fetch), "FETCH_SIZE")
    write = _per_kernel(_rows(args.write), "WRITE_SIZE")
    copy_name, copy_fetch = _pick(fetch, "k_copy")
    _, copy_write = _pick(write, "k_copy")
    copy_bytes = args.copy_bytes
    f_scale = copy_bytes / (copy_fetch * 1024.0)
    w_scale = copy_bytes / (copy_write * 1024.0)
    kname, kfetch = _pick(fetch, args.kernel, args.grid_size, args.first)
    _, kwrite = _pick(write, args.kernel, args.grid_size, args.first)
    read_b = kfetch * 1024.0 * f_scale
    write_b = kwrite * 1024.0 * w_scale
    algo = args.cells * args.bytes_per_lu
    rec = {
        "kernel": kname, "command": args.cmd, "grid_size": args.grid_size,
        "dispatches": len(_values(fetch, kname, args.grid_size, args.first)),
        "fetch_kib_per_dispatch": kfetch, "write_kib_per_dispatch": kwrite,
        "calibration": {"kernel": copy_name, "bytes_read_and_written": copy_bytes,
                        "fetch_kib": copy_fetch, "write_kib": copy_write,
                        "fetch_scale": f_scale, "write_scale": w_scale},
        "read_bytes_per_launch": read_b, "write_bytes_per_launch": write_b,
        "bytes_per_launch": read_b + write_b, "algorithmic_bytes_per_launch": algo,
        "traffic_over_algorithmic": (read_b + write_b) / algo,
    }
    Path(args.out).write_text(json.dumps(rec, indent=1) + "\n")


def span(args):
    """The last uninterrupted run of dispatches whose names contain one of the given
    substrings, cut where the GPU idled longer than --split-gap-us between two of them
    (a host synchronisation: e.g. the timed run of the batched engine, whose hipGraph
    replays only pass kernels back to back): dispatches per kernel, the span from the
    first start to the last end, the time the GPU was busy with them (union of
    intervals), and cells x steps / span."""
    rows = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                  for r in _rows(args.csv))
    match = [any(k in n for k in args.kernel) for _, _, n in rows]
    end = max((i for i, m in enumerate(match) if m), default=None)
    if end is None:
        raise SystemExit(f"no dispatch of {args.kernel}")
    start = end
    gap = args.split_gap_us * 1e3
    while start > 0 and match[start - 1] and rows[start][0] - rows[start - 1][1] <= gap:
        start -= 1
    sel = rows[start:end + 1]
    per = {}
    for t0, t1, n in sel:
        per.setdefault(n, []).append(t1 - t0)
    t_first, t_last = sel[0][0], max(t1 for _, t1, _ in sel)
    busy, cur0, cur1 = 0, None, None
    for t0, t1, _ in sel:
        if cur1 is None or t0 > cur1:
            busy += 0 if cur1 is None else cur1 - cur0
            cur0, cur1 = t0, t1
        else:
            cur1 = max(cur1, t1)
    busy += cur1 - cur0
    sp = t_last - t_first
    lines = [f"# rocprofv3 kernel trace: {args.title}", "", f"Command: `{args.cmd}`", "",
             f"The last uninterrupted run of dispatches of {', '.join(args.kernel)} (no idle "
             f"gap above {args.split_gap_us:g} us): {len(sel)} dispatches.", "",
             "| kernel | dispatches | avg us | total ms |", "|---|---:|---:|---:|"]
    for n, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        lines.append(f"| `{n[:60]}` | {len(d)} | {sum(d) / len(d) / 1e3:.1f} | {sum(d) / 1e6:.2f} |")
    lines += ["", f"Span of these dispatches: {sp / 1e6:.2f} ms, of which kernels busy "
                  f"{busy / 1e6:.2f} ms (gaps {(sp - busy) / 1e6:.2f} ms)."]
    if args.cells and args.steps:
        lines[-1] += (f"  {args.cells:.0f} cells x {args.steps} steps / span = "
                      f"{args.cells * args.steps / (sp * 1e-9) / 1e6:.0f} MLUPS.")
    Path(args.out).write_text("\n".join(lines) + "\n")


SQ_COUNTERS = ["SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD",
               "SQ_INSTS_VMEM_WR", "SQ_WAVES"]


def sq(args):
    """Instruction counts per dispatch (median over the kernel's dispatches with the given
    grid sizes) beside the launch time of the same dispatches in a kernel trace, and the
    share of the VALU issue capacity they take: SQ_INSTS_VALU x 4 cycles (a wave64 FP64 or
    unpacked FP32 instruction on a 16-lane SIMD) / (SIMDs x clock x launch time)."""
    crow = _rows(args.counters)
    trows = _rows(args.trace)
    grids = set(args.grid_size)

    def grid_ok(r, key):
        return not grids or int(r.get(key) or 0) in grids

    lines = [f"# rocprofv3 instruction counters (SQ), {args.title}", "",
             f"Counters: `{args.cmd}`", "", f"Launch times: `{args.trace_cmd}`", "",
             "Medians per dispatch over the kernel's dispatches"
             + (f" with grid size {', '.join(map(str, args.grid_size))}" if grids else "")
             + " (millions of wavefront instructions; waves in units).  VALU busy = "
             f"SQ_INSTS_VALU x 4 / ({args.simds} SIMDs x {args.clock_ghz} GHz x launch time).",
             "", "| kernel | VALU | SALU | LDS | VMEM rd | VMEM wr | waves | launch ms | VALU busy |",
             "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for spec in args.kernel:
        # "<substring>@N": only the kernel's first N dispatches (in each file)
        sub, _, first = spec.rpartition("@") if "@" in spec else (spec, "", "")
        first = int(first) if first else 0
        rows = [r for r in crow if sub in r["Kernel_Name"] and r["Counter_Name"] in SQ_COUNTERS
                and grid_ok(r, "Grid_Size")]
        if first:
            ids = sorted({int(r["Dispatch_Id"]) for r in rows})[:first]
            rows = [r for r in rows if int(r["Dispatch_Id"]) in set(ids)]
        vals = {}
        name = None
        for r in rows:
            vals.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
            name = r["Kernel_Name"]
        if name is None:
            raise SystemExit(f"no counter rows for {sub!r}")
        med = {k: statistics.median(v) for k, v in vals.items()}
        tsel = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"])) for r in trows
                      if sub in r["Kernel_Name"] and grid_ok(r, "Grid_Size_X"))
        if first:
            tsel = tsel[:first]
        durs = [e - b for b, e in tsel]
        if not durs:
            raise SystemExit(f"no trace dispatch of {sub!r}")
        t = statistics.median(durs) * 1e-9
        busy = med.get("SQ_INSTS_VALU", 0.0) * 4 / (args.simds * args.clock_ghz * 1e9 * t)
        m = lambda k: f"{med.get(k, float('nan')) / 1e6:.2f}"
        lines.append(f"| `{name[:46]}` | {m('SQ_INSTS_VALU')} | {m('SQ_INSTS_SALU')} | "
                     f"{m('SQ_INSTS_LDS')} | {m('SQ_INSTS_VMEM_RD')} | {m('SQ_INSTS_VMEM_WR')} | "
                     f"{med.get('SQ_WAVES', 0):.0f} | {t * 1e3:.4f} | {busy:.2f} |")
    Path(args.out).write_text("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("out")
    a.add_argument("--cmd", required=True)
    b = sub.add_parser("traffic")
    b.add_argument("fetch")
    b.add_argument("write")
    b.add_argument("out")
    b.add_argument("--kernel", required=True)
    b.add_argument("--cells", type=float, required=True)
    b.add_argument("--bytes-per-lu", type=float, required=True)
    b.add_argument("--copy-bytes", type=float, default=float(4 << 30))
    b.add_argument("--grid-size", type=int, default=None,
                   help="only dispatches of the kernel with this grid size (threads)")
    b.add_argument("--first", type=int, default=0,
                   help="only the kernel's first N dispatches (the copy probe: all)")
    b.add_argument("--cmd", required=True)
    c = sub.add_parser("trace")
    c.add_argument("csv")
    c.add_argument("out")
    c.add_argument("--kernel", action="append", required=True,
                   help="<name substring>=<algorithmic bytes per launch>, repeatable")
    c.add_argument("--last", type=int, default=50)
    c.add_argument("--first", type=int, default=0,
                   help="the first N dispatches of each kernel instead of the last")
    c.add_argument("--grid-size", type=int, action="append", default=[],
                   help="only dispatches with one of these grid sizes (threads)")
    c.add_argument("--cmd", required=True)
    w = sub.add_parser("window")
    w.add_argument("csv")
    w.add_argument("out")
    w.add_argument("--sequence", nargs="+", required=True,
                   help="kernel-name substrings of consecutive dispatches")
    w.add_argument("--bytes", nargs="*", type=float, default=[],
                   help="algorithmic bytes per launch of each dispatch in the sequence")
    w.add_argument("--skip", type=int, default=0, help="leading dispatches before the window")
    w.add_argument("--title", default="timed region")
    w.add_argument("--cmd", required=True)
    sp = sub.add_parser("span")
    sp.add_argument("csv")
    sp.add_argument("out")
    sp.add_argument("--kernel", action="append", required=True, help="name substring, repeatable")
    sp.add_argument("--cells", type=float, default=0.0)
    sp.add_argument("--steps", type=int, default=0)
    sp.add_argument("--split-gap-us", type=float, default=50.0)
    sp.add_argument("--title", default="dispatch span")
    sp.add_argument("--cmd", required=True)
    q = sub.add_parser("sq")
    q.add_argument("counters", help="counter_collection.csv of a --pmc run of SQ_ counters")
    q.add_argument("trace", help="kernel_trace.csv of a run of the same command")
    q.add_argument("out")
    q.add_argument("--kernel", action="append", required=True,
                   help="name substring, repeatable; <substring>@N for the first N dispatches")
    q.add_argument("--grid-size", type=int, action="append", default=[],
                   help="only dispatches with one of these grid sizes (threads)")
    q.add_argument("--simds", type=int, default=1024)
    q.add_argument("--clock-ghz", type=float, default=2.4)
    q.add_argument("--title", default="")
    q.add_argument("--cmd", required=True)
    q.add_argument("--trace-cmd", required=True)
    args = ap.parse_args()
    {"stats": stats, "traffic": traffic, "trace": trace, "window": window,
     "span": span, "sq": sq}[args.mode](args)


if __name__ == "__main__":
    main()
