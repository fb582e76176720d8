"""Profiler driver and counter summaries for the tick kernels.

Two uses:

``python -m polesitter_amd.profiling run --workload plummer1m --steps 3``
    The program to put under ``rocprofv3`` (kernel trace, or one ``--pmc``
    pass at a time).  It first copies ``--copy-mb`` MiB between two device
    buffers three times with a plain streaming kernel (``op_copy_kernel``,
    known byte counts: the calibration of the HBM byte counters), then runs
    ``--steps`` frame steps of the workload (tick + id-gather + Euler), one
    ``trav_init_kernel`` dispatch per tick.

``python -m polesitter_amd.profiling summarize --fetch DIR --write DIR``
    Reads rocprofv3's CSV output (``*_counter_collection.csv`` of a
    FETCH_SIZE pass and of a WRITE_SIZE pass) and prints a Markdown
    summary: per kernel family the counter
    values per tick, the copy kernel's counter reading against its known
    bytes (the scale factor), and the calibrated HBM bytes per tick.  The
    last line, ``traffic_bytes_per_tick <family> <bytes>``, is what bench.py
    reads from the committed summary.

``python -m polesitter_amd.profiling trace-summary --dir DIR --cmd "..."``
    Markdown table of a ``--kernel-trace --stats`` run, with per-tick sums of
    each kernel family.

``python -m polesitter_amd.profiling counters --dirs DIR... --command "..."``
    Markdown table of ``--pmc`` passes (any counters): each counter summed
    over every kernel family's dispatches of the run's last tick.

FETCH_SIZE and WRITE_SIZE are KiB per dispatch (derived TCC counters).
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict

import numpy as np

# kernel families: name fragments of the kernels that make up one stage
FAMILIES = {
    "trav_step": ("trav_step_kernel",),
    "p2p": ("p2p_kernel", "p2p_big_kernel", "p2p_collect_kernel"),
    "sort": ("sort_upsweep_kernel", "sort_rowscan_kernel", "sort_downsweep_kernel"),
    "build": ("tree_count_kernel", "tree_levels_kernel", "tree_write_kernel", "tree_mask_kernel"),
    "upward": ("upward_level_kernel", "upward_level_dev_kernel"),
    "l2l": ("leaf_locals_kernel", "l2l_level_kernel", "l2l_level_dev_kernel"),
    "prepare": ("bbox_partial_kernel", "bbox_final_kernel", "morton_kernel", "bbox_seg_kernel",
                "morton_seg_kernel"),
    "mlp": ("mlp_apply_kernel",),
    "frame": ("frame_kernel", "frame_finish_kernel"),
    "copy": ("op_copy_kernel",),
}
TICK_MARK = "trav_init_kernel"


def family_of(name: str):
    for fam, frags in FAMILIES.items():
        if any(f in name for f in frags):
            return fam
    return None


def run(args):
    from . import api
    from . import workloads as W

    L = api.hip_lib()
    # calibration copies
    nbytes = args.copy_mb << 20
    src = api.DeviceArray(nbytes // 4, np.float32)
    dst = api.DeviceArray(nbytes // 4, np.float32)
    for _ in range(3):
        api.check("ps_hip_op_copy", L.ps_hip_op_copy(src.ptr, dst.ptr, nbytes))
    print(f"profiling: copied {nbytes} bytes x3", flush=True)
    del src, dst

    if args.workload == "plummer1m":
        n = 1 << 20
        x, y, z, m = W.plummer(n, seed=1)
        v = [np.zeros(n, np.float32) for _ in range(3)]
    elif args.workload == "twoplummer16m":
        n = 1 << 24
        x, y, z, m, vx, vy, vz = W.two_plummers(n, seed=1)
        v = [vx, vy, vz]
    else:
        raise SystemExit(f"unknown workload {args.workload}")
    s = api.Solver(n, theta=args.theta, kahan=args.kahan)
    s.load(x, y, z, m)
    vel = [api.DeviceArray.from_numpy(vv) for vv in v]
    tmp = api.DeviceArray(3 * n, np.float32)
    for k in range(args.steps):  # the bench's step: tick, then the fused frame step
        s.tick(api.MORTON_AVX2)
        s.frame_step(vel, tmp, 1e-3)
        api.check("sync", L.ps_hip_synchronize())
        print(f"profiling: step {k + 1}/{args.steps} done", flush=True)
    s.close()
    return 0


def _csvs(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))


def read_counters(d, counter):
    """Per dispatch value of one counter: {dispatch_id: (kernel, value)}."""
    out = {}
    for path in _csvs(d, "counter_collection.csv"):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                key = (path, r["Dispatch_Id"])
                name = r["Kernel_Name"]
                val = float(r["Counter_Value"])
                if key in out:
                    out[key] = (name, out[key][1] + val)
                else:
                    out[key] = (name, val)
    return out


def per_family(disp):
    fam = defaultdict(float)
    ticks = 0
    copies = []
    for name, val in disp.values():
        if TICK_MARK in name:
            ticks += 1
        f = family_of(name)
        if f == "copy":
            copies.append(val)
        elif f:
            fam[f] += val
    return fam, ticks, copies


def summarize(args):
    copy_bytes = args.copy_mb << 20
    fetch = read_counters(args.fetch, "FETCH_SIZE")
    write = read_counters(args.write, "WRITE_SIZE")
    ff, tf, cf = per_family(fetch)
    fw, tw, cw = per_family(write)
    if not cf or not cw or not tf or not tw:
        print("no copy-kernel or tick dispatches found", file=sys.stderr)
        return 1
    # the copy kernel reads and writes copy_bytes each: counter KiB -> bytes
    kf = copy_bytes / (float(np.median(cf)) * 1024.0)
    kw = copy_bytes / (float(np.median(cw)) * 1024.0)
    lines = [
        f"# HBM traffic per tick ({args.workload})",
        "",
        "Commands (each pass its own run, from the repository root):",
        "",
        "```",
        f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d <dir_fetch> -o run -- "
        f"python3 -u -m polesitter_amd.profiling run --workload {args.workload} "
        f"--steps {args.steps} --copy-mb {args.copy_mb}",
        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d <dir_write> -o run -- "
        f"python3 -u -m polesitter_amd.profiling run --workload {args.workload} "
        f"--steps {args.steps} --copy-mb {args.copy_mb}",
        f"python3 -m polesitter_amd.profiling summarize --fetch <dir_fetch> --write <dir_write> "
        f"--workload {args.workload} --steps {args.steps} --copy-mb {args.copy_mb}",
        "```",
        "",
        f"Calibration: `op_copy_kernel` moves {copy_bytes} B each way per dispatch; "
        f"FETCH_SIZE reads {np.median(cf):.0f} KiB, WRITE_SIZE {np.median(cw):.0f} KiB "
        f"(median of {len(cf)}): scale factors fetch {kf:.4f}, write {kw:.4f} "
        f"(bytes per counter KiB / 1024).",
        "",
        f"Ticks profiled: {tf} (fetch pass), {tw} (write pass).",
        "",
        "| kernel family | FETCH_SIZE KiB/tick | WRITE_SIZE KiB/tick | calibrated HBM bytes/tick |",
        "|---|---|---|---|",
    ]
    traffic = {}
    for fam in FAMILIES:
        if fam == "copy" or (fam not in ff and fam not in fw):
            continue
        a = ff.get(fam, 0.0) / tf
        b = fw.get(fam, 0.0) / tw
        t = a * 1024.0 * kf + b * 1024.0 * kw
        traffic[fam] = t
        lines.append(f"| {fam} | {a:.0f} | {b:.0f} | {t:.4g} |")
    lines.append("")
    for fam, t in traffic.items():
        lines.append(f"traffic_bytes_per_tick {fam} {t:.6g}")
    print("\n".join(lines))
    return 0


def trace_summary(args):
    """Markdown summary of a --kernel-trace --stats run: the kernel table and
    per-tick sums of each kernel family (ticks = trav_init_kernel dispatches)."""
    stats = _csvs(args.dir, "kernel_stats.csv")
    trace = _csvs(args.dir, "kernel_trace.csv")
    if not stats or not trace:
        print("no kernel_stats.csv / kernel_trace.csv", file=sys.stderr)
        return 1
    with open(stats[0], newline="") as f:
        rows = list(csv.DictReader(f))
    fam_ns = defaultdict(float)
    fam_n = defaultdict(int)
    ticks = 0
    with open(trace[0], newline="") as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if TICK_MARK in name:
                ticks += 1
            fm = family_of(name)
            if fm:
                fam_ns[fm] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                fam_n[fm] += 1
    lines = [f"# Kernel trace: {args.title}", "", "Command (from the repository root):", "",
             "```", args.cmd, "```", "",
             f"Ticks in the run: {ticks} (one `trav_init_kernel` per tick).", "",
             "| kernel | calls | total ms | avg us | % |", "|---|---|---|---|---|"]
    for r in rows[:args.top]:
        name = r["Name"].replace("(anonymous namespace)::", "")
        name = name.split("(")[0] if not name.startswith("void ") else name[5:].split("(")[0]
        lines.append(f"| `{name}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.1f} |")
    if ticks:
        lines += ["", "Per tick, summed over each family's dispatches:", "",
                  "| family | dispatches / tick | ms / tick |", "|---|---|---|"]
        for fm in FAMILIES:
            if fm in fam_ns:
                lines.append(f"| {fm} | {fam_n[fm] / ticks:.1f} | {fam_ns[fm] / ticks / 1e6:.4f} |")
    print("\n".join(lines))
    return 0


def counters(args):
    """Markdown table of --pmc passes: every counter summed over each kernel
    family's dispatches of the last tick of the run (the dispatches after the
    last trav_init_kernel), plus derived issue figures for the families."""
    disp = defaultdict(dict)  # (path, dispatch) -> {counter: value}
    names, order = {}, {}
    for d in args.dirs:
        for path in _csvs(d, "counter_collection.csv"):
            with open(path, newline="") as f:
                for r in csv.DictReader(f):
                    key = (path, int(r["Dispatch_Id"]))
                    c = r["Counter_Name"]
                    disp[key][c] = disp[key].get(c, 0.0) + float(r["Counter_Value"])
                    names[key] = r["Kernel_Name"]
                    order[key] = int(r["Dispatch_Id"])
    fam = defaultdict(lambda: defaultdict(float))
    for path in sorted({k[0] for k in disp}):
        keys = sorted((k for k in disp if k[0] == path), key=lambda k: order[k])
        marks = [i for i, k in enumerate(keys) if TICK_MARK in names[k]]
        if not marks:
            continue
        for k in keys[marks[-1]:]:
            f = family_of(names[k])
            if f:
                for c, v in disp[k].items():
                    fam[f][c] += v
    lines = [f"# Counters: {args.title}", "", "Commands (one pass each, from the repository "
             "root):", "", "```"] + list(args.command) + ["```", "",
             "Summed over the last tick's dispatches of each kernel family.", ""]
    cols = sorted({c for f in fam.values() for c in f})
    fams = [f for f in FAMILIES if f in fam]
    lines += ["| counter | " + " | ".join(fams) + " |", "|---|" + "---|" * len(fams)]
    for c in cols:
        lines.append(f"| {c} | " + " | ".join(f"{fam[f].get(c, 0.0):.4g}" for f in fams) + " |")
    if args.ms:
        lines += ["", "Issue figures (VALU wave-instructions x 4 cycles over the SIMD cycles of "
                  f"{args.simds} SIMDs at {args.ghz} GHz in the family's time per tick):", ""]
        for spec in args.ms:
            f, ms = spec.split("=")
            v = fam[f].get("SQ_INSTS_VALU", 0.0)
            cyc = float(ms) * 1e-3 * args.ghz * 1e9 * args.simds
            lines.append(f"* {f}: {v:.4g} VALU wave-instructions in {float(ms):.3f} ms -> "
                         f"{100.0 * 4 * v / cyc:.1f} % of the SIMD issue cycles")
    print("\n".join(lines))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m polesitter_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--workload", default="plummer1m")
    r.add_argument("--steps", type=int, default=3)
    r.add_argument("--theta", type=float, default=2.0)
    r.add_argument("--kahan", action="store_true")
    r.add_argument("--copy-mb", type=int, default=512)
    m = sub.add_parser("summarize")
    m.add_argument("--fetch", required=True)
    m.add_argument("--write", required=True)
    m.add_argument("--workload", default="plummer1m")
    m.add_argument("--steps", type=int, default=3)
    m.add_argument("--copy-mb", type=int, default=512)
    t = sub.add_parser("trace-summary")
    t.add_argument("--dir", required=True)
    t.add_argument("--title", default="tick")
    t.add_argument("--cmd", default="")
    t.add_argument("--top", type=int, default=16)
    c = sub.add_parser("counters")
    c.add_argument("--dirs", nargs="+", required=True)
    c.add_argument("--title", default="tick")
    c.add_argument("--command", action="append", default=[])
    c.add_argument("--ms", action="append", default=[],
                   help="family=ms per tick, for the issue figures")
    c.add_argument("--simds", type=int, default=1024)
    c.add_argument("--ghz", type=float, default=2.4)
    a = ap.parse_args(argv)
    if a.cmd == "run":
        return run(a)
    if a.cmd == "summarize":
        return summarize(a)
    if a.cmd == "counters":
        return counters(a)
    return trace_summary(a)


if __name__ == "__main__":
    sys.exit(main())
