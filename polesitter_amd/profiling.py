"""Profiler driver and counter summaries for the tick kernels.

Two uses:

``python -m polesitter_amd.profiling run --workload plummer1m --steps 3``
    The program to put under ``rocprofv3`` (kernel trace, or one ``--pmc``
    pass at a time).  It first runs four streaming kernels of known traffic
    over buffers of ``--cal-mb`` MiB, three times each (larger than the
    256 MiB last-level MALL, so they stream from HBM): a read-only pass
    (``op_read_kernel``), a write-only pass (``op_fill_kernel``), a copy
    (``op_copy_kernel``, half the size each way) and a random 4-byte gather
    (``op_gather_kernel``).  Then it runs ``--steps`` frame steps of the
    workload (tick + id-gather + Euler), one ``trav_init_kernel`` dispatch
    per tick.

``python -m polesitter_amd.profiling summarize --dirs DIR... --trace DIR``
    Reads rocprofv3's CSV output of the four counter passes (FETCH_SIZE;
    WRITE_SIZE; the read requests by size TCC_EA0_RDREQ_32B/_64B/_128B and
    TCC_EA0_RDREQ; TCC_BUBBLE, TCC_EA0_RDREQ_DRAM, TCC_EA0_WRREQ and
    TCC_EA0_WRREQ_64B) and optionally a kernel trace of the same program.
    It checks each byte formula against the calibration kernels' known
    bytes, then prints per kernel family the bytes per tick and, with the
    trace, their rate.  The last lines, ``traffic_bytes_per_tick <family>
    <bytes>``, are what bench.py reads from the committed summary.

``python -m polesitter_amd.profiling trace-summary --dir DIR --cmd "..."``
    Markdown table of a ``--kernel-trace --stats`` run, with per-tick sums of
    each kernel family.

``python -m polesitter_amd.profiling counters --dirs DIR... --command "..."``
    Markdown table of ``--pmc`` passes (any counters): each counter summed
    over every kernel family's dispatches of the run's last tick.

FETCH_SIZE and WRITE_SIZE are KiB per dispatch (derived TCC counters);
the raw TCC counters are request counts per dispatch, summed over the
channels and XCDs.
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
    "sort": ("sort_upsweep_kernel", "sort_rowscan_kernel", "sort_downsweep_kernel",
             "delta_mask_kernel", "delta_scan_kernel", "delta_compact_kernel",
             "delta_pack_kernel", "delta_small_sort_kernel", "delta_merge_kernel",
             "delta_scatter_kernel", "copy6_kernel"),
    "build": ("tree_count_kernel", "tree_levels_kernel", "tree_write_kernel", "tree_mask_kernel"),
    "upward": ("upward_level_kernel", "upward_level_dev_kernel"),
    "l2l": ("leaf_locals_kernel", "l2l_level_kernel", "l2l_level_dev_kernel"),
    "prepare": ("bbox_partial_kernel", "bbox_final_kernel", "morton_kernel", "bbox_seg_kernel",
                "morton_seg_kernel"),
    "mlp": ("mlp_apply_kernel",),
    "frame": ("frame_kernel", "frame_finish_kernel"),
    "cal_copy": ("op_copy_kernel",),
    "cal_read": ("op_read_kernel",),
    "cal_fill": ("op_fill_kernel",),
    "cal_gather": ("op_gather_kernel",),
}
CAL = ("cal_read", "cal_fill", "cal_copy", "cal_gather")
TICK_MARK = "trav_init_kernel"


def family_of(name: str):
    for fam, frags in FAMILIES.items():
        if any(f in name for f in frags):
            return fam
    return None


def cal_bytes(cal_mb: int):
    """Known traffic of one dispatch of each calibration kernel: (read, write)."""
    nb = cal_mb << 20
    blocks = min(8192, (nb // 16 + 255) // 256)
    g = gather_elems(cal_mb)
    return {"cal_read": (nb, 16 * blocks), "cal_fill": (0, nb), "cal_copy": (nb // 2, nb // 2),
            # the gather reads its index array and one 4-byte element per
            # index (the request size decides the fetched bytes), writes 4 B
            "cal_gather": (4 * g + 4 * g, 4 * g)}


def gather_elems(cal_mb: int) -> int:
    return (cal_mb << 20) // 64


def run(args):
    from . import api
    from . import workloads as W

    L = api.hip_lib()
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

    def step():  # the bench's step: tick, then the fused frame step
        s.tick(api.MORTON_AVX2)
        s.frame_step(vel, tmp, 1e-3)
        api.check("sync", L.ps_hip_synchronize())

    # the first tick (a full sort; later ones sort incrementally) comes before
    # the calibration kernels, and the summary counts the ticks after them
    step()
    print("profiling: warm-up step done", flush=True)
    nb = args.cal_mb << 20
    a = api.DeviceArray(nb // 4, np.float32)
    b = api.DeviceArray(nb // 4, np.float32)
    g = gather_elems(args.cal_mb)
    rng = np.random.default_rng(5)
    idx = api.DeviceArray.from_numpy(rng.integers(0, nb // 4, g, dtype=np.uint32))
    out = api.DeviceArray(g, np.uint32)
    for _ in range(3):
        api.check("fill", L.ps_hip_op_bandwidth(2, None, a.ptr, None, nb // 16))
        api.check("read", L.ps_hip_op_bandwidth(1, a.ptr, b.ptr, None, nb // 16))
        api.check("copy", L.ps_hip_op_bandwidth(0, a.ptr, b.ptr, None, nb // 32))
        api.check("gather", L.ps_hip_op_bandwidth(3, a.ptr, out.ptr, idx.ptr, g))
    print(f"profiling: calibration kernels over {nb} bytes x3", flush=True)
    del a, b, idx, out
    for k in range(args.steps):
        step()
        print(f"profiling: step {k + 1}/{args.steps} done", flush=True)
    s.close()
    return 0


def _csvs(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))


def _after_warmup(rows):
    """rows: [(dispatch id, kernel name, row)] of one file.  The tick
    dispatches before the first calibration kernel (the warm-up step of a
    `run`) are dropped; a file without calibration kernels is kept whole."""
    first_cal = min((d for d, name, _ in rows if family_of(name) in CAL), default=None)
    if first_cal is None:
        return rows
    return [(d, name, r) for d, name, r in rows
            if d >= first_cal or family_of(name) in CAL]


def read_all(dirs):
    """{(path, dispatch): (kernel, {counter: value})} over the passes in dirs
    (without a run's warm-up step)."""
    out = {}
    for d in dirs:
        for path in _csvs(d, "counter_collection.csv"):
            with open(path, newline="") as f:
                rows = [(int(r["Dispatch_Id"]), r["Kernel_Name"], r) for r in csv.DictReader(f)]
                for _, _, r in _after_warmup(rows):
                    key = (path, int(r["Dispatch_Id"]))
                    name, vals = out.get(key, (r["Kernel_Name"], {}))
                    c = r["Counter_Name"]
                    vals[c] = vals.get(c, 0.0) + float(r["Counter_Value"])
                    out[key] = (name, vals)
    return out


def _formulas(v):
    """Byte estimates of one dispatch from whichever counters it carries."""
    out = {}
    if "FETCH_SIZE" in v:
        out["fetch_size"] = 1024.0 * v["FETCH_SIZE"]
    if all(k in v for k in ("TCC_EA0_RDREQ_32B", "TCC_EA0_RDREQ_64B", "TCC_EA0_RDREQ_128B")):
        out["read_by_size"] = (32.0 * v["TCC_EA0_RDREQ_32B"] + 64.0 * v["TCC_EA0_RDREQ_64B"] +
                               128.0 * v["TCC_EA0_RDREQ_128B"])
    if "WRITE_SIZE" in v:
        out["write_size"] = 1024.0 * v["WRITE_SIZE"]
    if "TCC_EA0_WRREQ" in v and "TCC_EA0_WRREQ_64B" in v:
        out["write_by_size"] = (32.0 * (v["TCC_EA0_WRREQ"] - v["TCC_EA0_WRREQ_64B"]) +
                                64.0 * v["TCC_EA0_WRREQ_64B"])
    return out


def family_ms_per_tick(trace_dir):
    """{family: ms per tick} from a --kernel-trace run of the same program."""
    fam_ns = defaultdict(float)
    ticks = 0
    for path in _csvs(trace_dir, "kernel_trace.csv"):
        with open(path, newline="") as f:
            rows = [(int(r["Dispatch_Id"]), r["Kernel_Name"], r) for r in csv.DictReader(f)]
            for _, _, r in _after_warmup(rows):
                name = r["Kernel_Name"]
                if TICK_MARK in name:
                    ticks += 1
                fm = family_of(name)
                if fm:
                    fam_ns[fm] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    if not ticks:
        return {}, 0
    out = {f: ns / 1e6 / (1 if f in CAL else ticks) for f, ns in fam_ns.items()}
    return out, ticks


def summarize(args):
    disp = read_all(args.dirs)
    # per family: summed formulas over the run (tick families) or the median
    # per dispatch (calibration kernels)
    fam = defaultdict(lambda: defaultdict(float))
    cal = defaultdict(lambda: defaultdict(list))
    ticks = defaultdict(int)  # per pass file: trav_init dispatches
    for (path, _), (name, vals) in disp.items():
        if TICK_MARK in name:
            ticks[path] += 1
        f = family_of(name)
        if not f:
            continue
        for k, b in _formulas(vals).items():
            if f in CAL:
                cal[f][k].append(b)
            else:
                fam[f][k] += b
        for c in ("TCC_EA0_RDREQ", "TCC_EA0_RDREQ_32B", "TCC_EA0_RDREQ_64B",
                  "TCC_EA0_RDREQ_128B", "TCC_BUBBLE", "TCC_EA0_RDREQ_DRAM"):
            if c in vals and f in CAL:
                cal[f][c].append(vals[c])
    n_ticks = max(ticks.values()) if ticks else 0
    if not n_ticks:
        print("no tick dispatches found", file=sys.stderr)
        return 1
    known = cal_bytes(args.cal_mb)
    med = lambda f, k: float(np.median(cal[f][k])) if cal[f].get(k) else float("nan")
    # scale factors: known bytes over what each formula reads for the
    # read-only and the write-only kernel (both far larger than the MALL)
    kr_fetch = known["cal_read"][0] / med("cal_read", "fetch_size")
    kr_size = known["cal_read"][0] / med("cal_read", "read_by_size")
    kw_write = known["cal_fill"][1] / med("cal_fill", "write_size")
    ms, _ = family_ms_per_tick(args.trace) if args.trace else ({}, 0)
    cmd = (f"python3 -u -m polesitter_amd.profiling run --workload {args.workload} "
           f"--steps {args.steps} --cal-mb {args.cal_mb}")
    lines = [
        f"# HBM traffic per tick ({args.workload})", "",
        "Commands (each pass its own run, from the repository root):", "", "```",
        f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d <d1> -o run -- {cmd}",
        f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d <d2> -o run -- {cmd}",
        f"rocprofv3 --pmc TCC_EA0_RDREQ_32B TCC_EA0_RDREQ_64B TCC_EA0_RDREQ_128B TCC_EA0_RDREQ "
        f"--output-format csv -d <d3> -o run -- {cmd}",
        f"rocprofv3 --pmc TCC_BUBBLE TCC_EA0_RDREQ_DRAM TCC_EA0_WRREQ TCC_EA0_WRREQ_64B "
        f"--output-format csv -d <d4> -o run -- {cmd}",
        f"rocprofv3 --kernel-trace --stats --output-format csv -d <d5> -o run -- {cmd}",
        f"python3 -m polesitter_amd.profiling summarize --dirs <d1> <d2> <d3> <d4> --trace <d5> "
        f"--workload {args.workload} --steps {args.steps} --cal-mb {args.cal_mb}",
        "```", "",
        "## Calibration (median per dispatch; buffers of "
        f"{args.cal_mb} MiB, the MALL is 256 MiB)", "",
        "| kernel | known read B | FETCH_SIZE B | read by size B | known write B | "
        "WRITE_SIZE B | RDREQ | 32B | 64B | 128B | BUBBLE | ms |",
        "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for f in CAL:
        if f not in cal:
            continue
        kr, kw = known[f]
        t = ms.get(f)
        lines.append(f"| {f} | {kr:.4g} | {med(f, 'fetch_size'):.4g} | "
                     f"{med(f, 'read_by_size'):.4g} | {kw:.4g} | {med(f, 'write_size'):.4g} | "
                     f"{med(f, 'TCC_EA0_RDREQ'):.4g} | {med(f, 'TCC_EA0_RDREQ_32B'):.4g} | "
                     f"{med(f, 'TCC_EA0_RDREQ_64B'):.4g} | {med(f, 'TCC_EA0_RDREQ_128B'):.4g} | "
                     f"{med(f, 'TCC_BUBBLE'):.4g} | "
                     f"{'' if t is None else f'{t / 3:.3f}'} |")
    lines += ["", f"Scale factors (known bytes / formula, read-only and write-only kernels): "
              f"FETCH_SIZE {kr_fetch:.4f}, read by size {kr_size:.4f}, WRITE_SIZE "
              f"{kw_write:.4f}.", ""]
    g = cal.get("cal_gather")
    if g and g.get("read_by_size"):
        gb = med("cal_gather", "read_by_size")
        gf = med("cal_gather", "fetch_size") * kr_fetch
        lines += [f"Gather check (scattered 4-byte reads, 4 B index + 4 B element per "
                  f"lookup): reads by size x factor = {gb * kr_size:.4g} B, FETCH_SIZE x "
                  f"factor = {gf:.4g} B, for {known['cal_gather'][0]:.4g} B of useful reads "
                  f"({(gb * kr_size) / (gather_elems(args.cal_mb)):.1f} B fetched per lookup).",
                  ""]
    lines += [f"Ticks profiled per pass: {n_ticks}.", "",
              "## Per tick and kernel family", "",
              "Reads = read requests by size x factor; writes = WRITE_SIZE x factor.", "",
              "| family | reads B/tick | (FETCH_SIZE x factor) | writes B/tick | total B/tick | "
              "ms/tick | rate GB/s |", "|---|---|---|---|---|---|---|"]
    traffic = {}
    for f in FAMILIES:
        if f in CAL or f not in fam:
            continue
        v = fam[f]
        rd = v.get("read_by_size", 0.0) * kr_size / n_ticks
        fs = v.get("fetch_size", 0.0) * kr_fetch / n_ticks
        wr = v.get("write_size", 0.0) * kw_write / n_ticks
        tot = rd + wr
        traffic[f] = tot
        t = ms.get(f)
        rate = "" if not t else f"{tot / (t * 1e-3) / 1e9:.0f}"
        lines.append(f"| {f} | {rd:.4g} | ({fs:.4g}) | {wr:.4g} | {tot:.4g} | "
                     f"{'' if t is None else f'{t:.4f}'} | {rate} |")
    lines.append("")
    for f, t in traffic.items():
        lines.append(f"traffic_bytes_per_tick {f} {t:.6g}")
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
                rows = [(int(r["Dispatch_Id"]), r["Kernel_Name"], r) for r in csv.DictReader(f)]
                for _, _, r in _after_warmup(rows):
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
    r.add_argument("--cal-mb", type=int, default=2048)
    m = sub.add_parser("summarize")
    m.add_argument("--dirs", nargs="+", required=True)
    m.add_argument("--trace", default="")
    m.add_argument("--workload", default="plummer1m")
    m.add_argument("--steps", type=int, default=3)
    m.add_argument("--cal-mb", type=int, default=2048)
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
