#!/usr/bin/env python3
"""FMM tick throughput on one MI355X (particles per second per tick).

Workload (default ``plummer1m``): N = 1,048,576 Plummer sphere (a = 1,
truncated at 20a, isotropic, equal masses 1/N, float32 SoA), theta = 2.0,
the reference's threaded-build traversal.  One step = one tick
(ps_prepare_particles + ps_calc_forces, particles device-resident) followed
by the caller's frame step on the device (id-gather of velocities,
semi-implicit Euler dt = 1e-3, ids reset for the next frame; one fused call),
so every tick re-sorts and rebuilds.

    python bench.py [--steps K] [--warmup W]
                    [--workload plummer1m|twoplummer16m|plummer4m|ensemble512|mlp4m|sweep]

Prints one JSON line.  `value` = N * K / (wall time of K steps), bracketed
by device synchronisation.  `roofline` is the dominant kernel's algorithmic
work per launch over its launch time, both from HIP events recorded on the
engine's stream inside the timed steps; `roofline.step` / `tick_frac` put
the whole step against the sum of its stages' rooflines.  `cpu_baseline`
times the reference's own PS_MULTITHREADING CPU path (oracle/_ref,
shipped-style flags) on this host at 1 thread, at the physical cores the
process may use and at every usable logical CPU (the host's facts are in
`cpu_baseline.host`), on a bounded sample of the same workload.  `fp64_*`
report the force error against an FP64 direct sum on 4096 seeded targets
matched through id[] (tests/accuracy.c:505-531), for the GPU tick and for
the reference on the same state.  The default run (plummer1m) adds the
metric's second size, N = 16,777,216 two Plummers, as `sizes.twoplummer16m`
with the same fields.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FMM tick particles/sec at N=1M,16M @ fixed RMS force err vs FP64 direct"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOP_P2P = 20    # per particle pair (GPU Gems 3 ch. 31 convention)
FLOP_M2L = 41    # per M2L pair, op count of the reference's M2L


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from polesitter_amd import workloads as W
    if name == "plummer1m":
        n = 1 << 20
        x, y, z, m = W.plummer(n, seed=1)
        v = [np.zeros(n, np.float32) for _ in range(3)]
        desc = "Plummer a=1 r<=20a, N=1048576, m=1/N, v0=0"
    elif name == "plummer4m":
        n = 1 << 22
        x, y, z, m = W.plummer(n, seed=1)
        v = [np.zeros(n, np.float32) for _ in range(3)]
        desc = "Plummer a=1 r<=20a, N=4194304, m=1/N, v0=0"
    elif name == "twoplummer16m":
        n = 1 << 24
        x, y, z, m, vx, vy, vz = W.two_plummers(n, seed=1)
        v = [vx, vy, vz]
        desc = "two Plummers a=1 at (+-5,0,0), v=(-+0.3,0,0), N=16777216, m=1/N"
    else:
        raise SystemExit(f"unknown workload {name}")
    return n, x, y, z, m, v, desc


def measured_traffic(workload, family):
    """Calibrated HBM bytes per tick of a kernel family from the committed
    rocprofv3 --pmc summary (profiles/traffic_<workload>.md), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    try:
        with open(path) as f:
            for line in f:
                t = line.split()
                if len(t) == 3 and t[0] == "traffic_bytes_per_tick" and t[1] == family:
                    return float(t[2])
    except OSError:
        pass
    return None


def fp64_error(n, pos_sorted_forces, ids, ref_fn, seed=4096):
    """RMS / max relative error on 4096 seeded original indices."""
    from polesitter_amd import workloads as W
    tg = W.sample_targets(n, 4096, seed=seed)
    inv = np.empty(n, np.int64)
    inv[ids] = np.arange(n)
    pos = inv[tg]
    ref = ref_fn(pos)
    f = pos_sorted_forces[pos].astype(np.float64)
    d = f - ref
    e2 = (d * d).sum(1)
    r2 = (ref * ref).sum(1)
    ok = r2 > 1e-24
    return float(np.sqrt(e2.sum() / r2.sum())), float(np.sqrt(e2[ok] / r2[ok]).max())


def host_cpus():
    """What this host gives the process: logical CPUs, the affinity set, a
    cgroup CPU quota, the physical cores inside the affinity set and the CPU
    model.  `usable` = the affinity set, capped by the quota."""
    info = dict(nproc=os.cpu_count())
    aff = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else \
        list(range(os.cpu_count() or 1))
    info["affinity"] = len(aff)
    quota = None
    try:
        q = open("/sys/fs/cgroup/cpu.max").read().split()
        if q and q[0] != "max":
            quota = float(q[0]) / float(q[1])
    except (OSError, ValueError, IndexError):
        pass
    info["cgroup_cpu_quota"] = quota
    cores = set()
    for c in aff:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            cores.add((open(base + "physical_package_id").read().strip(),
                       open(base + "core_id").read().strip()))
        except OSError:
            pass
    info["physical_cores_in_affinity"] = len(cores) or None
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info["model"] = model
    usable = len(aff)
    if quota:
        usable = max(1, min(usable, int(quota)))
    info["usable"] = usable
    return info


def cpu_thread_counts(info):
    """Thread counts the reference is timed at: 1, the physical cores the
    process may use, and every usable logical CPU."""
    usable = info["usable"]
    phys = info["physical_cores_in_affinity"] or usable
    return sorted({1, max(1, min(phys, usable)), usable})


def cpu_baseline(n, x, y, z, m, theta, ticks, gpu_ref_fn, threads, info, err_threads=None):
    """The reference CPU path (oracle/_ref, shipped-style flags,
    PS_MULTITHREADING) on this host: one warm-up and `ticks` timed ticks per
    thread count, each from the workload's initial state; value = the
    fastest thread count's median.  The FP64 error of the reference on the
    same state is taken at err_threads (default: the fastest count)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if not O.ref_available("shipped_mt"):
        # no compiled reference: the single-threaded C restatement
        t0 = time.time()
        O.tick(x, y, z, m, theta=theta)
        tick_s = time.time() - t0
        return dict(value=n / tick_s, unit="particles/s", cores=1, kind="port",
                    sample="1 tick of the full workload, oracle/ps_oracle.c, 1 thread",
                    tick_s=tick_s, host=info)
    per = {}
    t_all = time.time()
    for thr in threads:
        # the reference caps a context at PS_MAX_THRDS + 1 = 33 threads
        # (H:384-386, H:1709-1714); a wider build lifts the cap
        build = "shipped_mt" if thr <= 33 else "shipped_mt_wide"
        if not O.ref_available(build):
            continue
        secs = []
        for k in range(ticks + 1):
            secs.append(float(O.ref_time_ticks(x, y, z, m, theta=theta, thrd_cnt=thr,
                                               n_ticks=1, build=build)[0]))
            log(f"bench: reference CPU tick {k + 1}/{ticks + 1} on {thr} threads: "
                f"{secs[-1]:.3f} s")
        per[thr] = dict(tick_s=float(np.median(secs[1:])), build=build)
    best = min(per, key=lambda t: per[t]["tick_s"])
    et = err_threads if err_threads in per else best
    r = O.ref_tick(x, y, z, m, theta=theta, thrd_cnt=et, build=per[et]["build"], dump=False)
    log(f"bench: reference CPU tick for the FP64 error done ({et} threads)")
    f = np.stack([r["fx"], r["fy"], r["fz"]], 1)
    err = fp64_error(n, f, r["id"], lambda pos: gpu_ref_fn(r, pos))
    wall = time.time() - t_all
    sample = (f"{ticks} ticks (+1 warm-up) of the full workload from its initial state per "
              f"thread count, median; src/polesitter.h built -O3 -ffast-math -march=x86-64-v3 "
              f"-DPS_MULTITHREADING (PS_MAX_THRDS=255 above 33 threads); thread counts "
              f"{sorted(per)} on {info['model']}; {wall:.1f} s wall")
    return dict(value=n / per[best]["tick_s"], unit="particles/s", cores=best, kind="reference",
                sample=sample, tick_s=per[best]["tick_s"],
                by_threads={str(t): dict(tick_s=round(v["tick_s"], 5),
                                         particles_per_s=round(n / v["tick_s"], 1))
                            for t, v in sorted(per.items())},
                fp64_rms_err=err[0], fp64_max_err=err[1], fp64_err_threads=et, host=info)


def _timed_steps(s, n, v, steps, warmup, mmode):
    """wall seconds per step (tick + id-gather + Euler) on a loaded Solver"""
    from polesitter_amd import api
    L = api.hip_lib()
    vel = [api.DeviceArray.from_numpy(vv) for vv in v]
    tmp = api.DeviceArray(3 * n, np.float32)

    def step():  # ids are 0..n-1 on entry (load, or the previous frame step)
        s.tick(mmode)
        s.frame_step(vel, tmp, 1e-3)
    for _ in range(warmup):
        step()
    api.check("sync", L.ps_hip_synchronize())
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    api.check("sync", L.ps_hip_synchronize())
    return (time.perf_counter() - t0) / steps


def bench_mlp(a):
    """The learned far-field correction at N = 4,194,304 Plummer (a = 1),
    judged against the best the plain FMM can do at the same accuracy.

    1. Plain settings on the evaluation realisation (seed 3): theta in
       --mlp-thetas, leaf capacity in --mlp-caps; FP64 RMS error on 4096
       targets and throughput per step (tick + id-gather + Euler).  The
       reference setting is theta --theta, leaf capacity 1; the baseline to
       beat is the fastest plain setting whose error does not exceed the
       reference setting's by more than --mlp-err-tol (1 %).
    2. The --mlp-candidates fastest plain settings that are faster than that
       baseline but less accurate get a correction each: trained on
       --mlp-ticks ticks of residuals of a training realisation (seed 1),
       validated on another (seed 2; the plain FMM judges, the correction is
       used only if it lowers the error), then measured on seed 3 like the
       plain settings.
    value = throughput of the fastest setting, plain or corrected, whose
    error does not exceed the reference setting's;
    speedup_vs_best_plain_at_equal_error says what the correction itself
    buys (1.0 when no corrected setting wins)."""
    from polesitter_amd import api, mlp
    from polesitter_amd import workloads as W
    n = 1 << 22
    mmode = api.MORTON_AVX2 if a.morton == "avx2" else api.MORTON_TRUNC
    L = api.hip_lib()
    zero = [np.zeros(n, np.float32)] * 3
    ex, ey, ez, em = W.plummer(n, seed=3)
    log(f"bench: {api.device_name()} workload mlp4m: Plummer N={n}")

    def measure(th, cap, params=None):
        with api.Solver(n, theta=th) as s:
            api.check("opt", api.ps_set_option(s.ctx, api.OPT_LEAF_CAP, cap))
            if params is not None:
                api.check("set_mlp", L.ps_hip_engine_set_mlp(s.ctx.engine, params.ctypes.data))
            s.load(ex, ey, ez, em)
            s.reset_ids()
            s.tick(mmode)
            r = s.result()
            f = np.stack([r["fx"], r["fy"], r["fz"]], 1)
            err = fp64_error(n, f, r["id"], s.direct_f64)
            del r, f
            s.load(ex, ey, ez, em)
            dt = _timed_steps(s, n, zero, a.steps, a.warmup, mmode)
            api.check("opt", api.ps_set_option(s.ctx, api.OPT_STATS, 2))
            s.reset_ids()
            s.tick(mmode)
            st = s.stats()
        return dict(theta=th, leaf_cap=cap, correction=params is not None, fp64_rms_err=err[0],
                    fp64_max_err=err[1], ms_per_step=round(1e3 * dt, 4),
                    particles_per_s=round(n / dt, 1),
                    stages_ms={k: round(st["ms_" + k], 4) for k in
                               ("sort", "build", "upward", "traverse", "l2l", "p2p", "mlp",
                                "total")})

    thetas = [float(t) for t in a.mlp_thetas.split(",")]
    caps = [int(c) for c in a.mlp_caps.split(",")]
    plain = []
    for th in thetas:
        for cap in caps:
            r = measure(th, cap)
            plain.append(r)
            log(f"bench: plain theta={th} cap={cap}: rms {r['fp64_rms_err']:.5f} "
                f"{r['ms_per_step']:.3f} ms")
    ref = next((r for r in plain if r["theta"] == a.theta and r["leaf_cap"] == 1), None)
    if ref is None:
        ref = measure(a.theta, 1)
        plain.append(ref)
    # "equal error": within --mlp-err-tol of the reference setting's (leaf
    # capacity alone moves it by 0.01 %; a strict comparison would crown
    # the reference setting itself as the plain baseline)
    bar = ref["fp64_rms_err"] * (1.0 + a.mlp_err_tol)
    ok_plain = [r for r in plain if r["fp64_rms_err"] <= bar]
    best_plain = max(ok_plain, key=lambda r: r["particles_per_s"])
    # candidates for a correction: faster than the best accurate plain
    # setting, too inaccurate on their own
    cands = sorted((r for r in plain if r["fp64_rms_err"] > bar and
                    r["particles_per_s"] > best_plain["particles_per_s"]),
                   key=lambda r: -r["particles_per_s"])[:a.mlp_candidates]
    corrected = []
    x, y, z, m = W.plummer(n, seed=1)
    val = W.plummer(n, seed=2) + (zero,)
    for c in cands:
        t0 = time.time()
        res = mlp.build_correction(x, y, z, m, zero, train_ticks=a.mlp_ticks, val_ticks=10,
                                   samples=4096, theta=c["theta"], leaf_cap=c["leaf_cap"],
                                   steps=3000, batch=8192, lr=2e-3, seed=1, val_system=val,
                                   log=log)
        train_s = time.time() - t0
        if not res["enabled"]:
            corrected.append(dict(c, correction=False, validation=dict(
                plain=res["val_plain"], corrected=res["val_corrected"], enabled=False)))
            continue
        r = measure(c["theta"], c["leaf_cap"], res["params"])
        r["validation"] = dict(plain=res["val_plain"], corrected=res["val_corrected"],
                               enabled=True, train_samples=res["train_samples"],
                               train_seconds=round(train_s, 1))
        r["plain_fp64_rms_err"] = c["fp64_rms_err"]
        corrected.append(r)
        log(f"bench: corrected theta={c['theta']} cap={c['leaf_cap']}: rms "
            f"{r['fp64_rms_err']:.5f} (plain {c['fp64_rms_err']:.5f}) {r['ms_per_step']:.3f} ms")
    ok_corr = [r for r in corrected if r.get("correction") and r["fp64_rms_err"] <= bar]
    best_corr = max(ok_corr, key=lambda r: r["particles_per_s"]) if ok_corr else None
    head = best_corr if (best_corr and best_corr["particles_per_s"] >
                         best_plain["particles_per_s"]) else best_plain
    line = dict(metric=METRIC, value=head["particles_per_s"], unit="particles/s", n_gpus=1,
                steps=a.steps, warmup=a.warmup, ms_per_step=head["ms_per_step"],
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic",
                config=dict(workload="mlp4m", n_particles=n, theta=a.theta, morton=a.morton,
                            thetas=thetas, caps=caps, train_ticks=a.mlp_ticks,
                            equal_error_tolerance=a.mlp_err_tol,
                            step="tick + id-gather + semi-implicit Euler dt=1e-3",
                            net="24-64-64-64-3 GELU (tanh form), FP32 weights; in-tick "
                                "on MFMA with bf16 hi/lo operand pairs, FP32 accumulation"),
                headline=dict(theta=head["theta"], leaf_cap=head["leaf_cap"],
                              correction=head["correction"]),
                reference_setting=ref, best_plain_at_equal_error=best_plain,
                best_corrected_at_equal_error=best_corr,
                speedup_vs_best_plain_at_equal_error=round(
                    (best_corr["particles_per_s"] if best_corr else 0.0) /
                    best_plain["particles_per_s"], 4) if best_corr else 1.0,
                speedup_vs_reference_setting=round(head["particles_per_s"] /
                                                   ref["particles_per_s"], 4),
                plain=plain, corrected=corrected)
    print(json.dumps(line), flush=True)
    return 0


def bench_ensemble(a):
    """512 independent systems (N_s log-uniform in [2048, 262144], even ones
    uniform cubes, odd ones Plummer with a ~ U[0.5, 2], masses 1/N_s) in one
    batched tick, recorded as a HIP graph and replayed; settings from the
    autotuner (leaf capacity per system-size class, P2P tile and lanes per
    batch) at each class's reference-setting FP64 error.  One step = one batched
    tick (prepare + calc_forces of every system); value = total particles /
    tick."""
    from polesitter_amd import api, autotune
    from polesitter_amd import workloads as W
    systems = W.ensemble(512, seed=1)
    n = sum(len(t[0]) for t in systems)
    mmode = api.MORTON_AVX2 if a.morton == "avx2" else api.MORTON_TRUNC
    log(f"bench: {api.device_name()} workload ensemble512: 512 systems, {n} particles")
    if a.no_tune:  # the reference setting, for profiling runs
        cfg = dict(caps=[1, 1, 1, 1], p2p_big=24, trav_lanes=8)
        tuned = dict(config=cfg, system_caps=[1] * len(systems), reference=None, ms_tick=None,
                     fp64_rms_err=None, class_errors=None, table=[], classes=None)
    else:
        tuned = autotune.tune_batch(systems, theta=a.theta, mode=mmode, log=log)
        cfg = tuned["config"]
    L = api.hip_lib()
    with api.Batch(n, len(systems), theta=a.theta, graph=True, kahan=a.kahan) as b:
        b.set_leaf_caps(tuned["system_caps"])
        b.set_option(api.OPT_P2P_BIG, cfg["p2p_big"])
        b.set_option(api.OPT_TRAV_LANES, cfg["trav_lanes"])
        if a.trav_blocks:
            b.set_option(api.OPT_TRAV_BLOCKS, a.trav_blocks)
        if a.trav_segments:
            b.set_option(api.OPT_TRAV_SEGMENTS, a.trav_segments)
        b.load(systems)
        a.warmup = max(a.warmup, 1)
        for _ in range(a.warmup):  # the first one records the graph
            b.tick(mmode)
        api.check("sync", L.ps_hip_synchronize())
        ms_dev = []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            b.tick(mmode)
            ms_dev.append(b.stats()["ms_tick"])
        api.check("sync", L.ps_hip_synchronize())
        wall = time.perf_counter() - t0
        st = b.stats()
    value = n * a.steps / wall
    line = dict(metric=METRIC, value=round(value, 1), unit="particles/s", n_gpus=1,
                steps=a.steps, warmup=a.warmup, ms_per_step=round(1e3 * wall / a.steps, 4),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic",
                config=dict(workload="ensemble512", n_systems=len(systems), n_particles=n,
                            theta=a.theta, morton=a.morton, kahan=bool(a.kahan),
                            traversal="MT depth-2 jobs (PS_MULTITHREADING)",
                            step="one batched tick (HIP graph replay) of all systems",
                            tuned=cfg, pointers="device-resident"),
                device_ms_per_tick=round(float(np.median(ms_dev)), 4),
                graph_replays=int(st["graph_replays"]), used_graph=int(st["used_graph"]),
                counts=dict(nodes=int(st["n_nodes"]), m2l_pairs=int(st["m2l_pairs"])),
                autotune=dict(reference=tuned["reference"], chosen_ms=tuned["ms_tick"],
                              chosen_fp64_rms_err=tuned["fp64_rms_err"],
                              chosen_class_errors=tuned["class_errors"],
                              classes=tuned["classes"],
                              speedup_vs_reference_setting=(
                                  None if a.no_tune else
                                  round(tuned["reference"]["ms_tick"] / tuned["ms_tick"], 4)),
                              candidates=[dict(c, ms=round(ms_, 3), err=round(e_, 6),
                                               class_errors={str(k): round(v, 6)
                                                             for k, v in per.items()})
                                          for c, ms_, per, e_ in tuned["table"]],
                              skipped=bool(a.no_tune)))
    print(json.dumps(line), flush=True)
    return 0


# tests/benchmark.c:136-137, and the reference's published times for them
# (seconds; docs/performance_graph_log.svg, hardware not stated): direct
# FP32 O(N^2), FMM thrd_cnt=1, FMM thrd_cnt=8, all at theta = 1.0
SWEEP_SIZES = (1000, 2500, 5000, 10000, 15000, 20000, 25000, 50000, 75000, 100000, 150000,
               200000)
PUBLISHED_SWEEP = {
    1000: (0.000315, 0.000459, 0.000341), 2500: (0.002507, 0.003155, 0.000596),
    5000: (0.007735, 0.003078, 0.003785), 10000: (0.02789, 0.00425, 0.004522),
    15000: (0.06154, 0.006603, 0.003075), 20000: (0.1097, 0.008098, 0.00389),
    25000: (0.1719, 0.01012, 0.005192), 50000: (0.6908, 0.01958, 0.01788),
    75000: (1.567, 0.02886, 0.01484), 100000: (2.866, 0.03768, 0.01975),
    150000: (6.924, 0.05461, 0.03162), 200000: (11.95, 0.07593, 0.04226)}


def bench_sweep(a):
    """tests/benchmark.c's sweep through the drop-in: N from 1,000 to
    200,000 in a uniform cube [0,100)^3 (glibc rand(), seed 7331), mass 1,
    the threaded build's traversal, at theta = 1.0 (the benchmark's) and 2.0.
    Per point: prepare + calc_forces on device arrays (median of --steps) and
    on host arrays as the benchmark passes them (median of 3), the
    reference's CPU path at thrd_cnt = 1, 8 and every usable CPU (median of
    3, after one warm-up), the published times, and the FP64 error of the
    GPU and of the reference.  At theta = 1.0 the reference's depth-2
    traversal roots make a target two levels below a containing source
    'far' (dist^2 = 1.6875 hw0^2 > 1.5625 hw0^2 for the corner depth-2 cells
    against the root), so those cells take the root's M2L and skip their
    near field: the GPU reproduces the same interaction lists, and the FP64
    error shows it.  value = GPU particles/s at N = 200,000, theta = 1.0."""
    from polesitter_amd import api
    from polesitter_amd import workloads as W
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    L = api.hip_lib()
    info = host_cpus()
    cpu_thr = sorted({1, 8, info["usable"]})
    mmode = api.MORTON_AVX2 if a.morton == "avx2" else api.MORTON_TRUNC
    rows = []
    log(f"bench: {api.device_name()} workload sweep: benchmark.c sizes, theta 1.0 and 2.0")
    for theta in (1.0, 2.0):
        for n in SWEEP_SIZES:
            x, y, z, m = W.benchmark_cube(n)
            row = dict(n=n, theta=theta)
            with api.Solver(n, theta=theta, stats=True) as s:
                s.load(x, y, z, m)
                s.tick(mmode)
                st = s.stats()
                r = s.result()
                f = np.stack([r["fx"], r["fy"], r["fz"]], 1)
                row["gpu_fp64_rms_err"] = fp64_error(n, f, r["id"], s.direct_f64)[0]
                row["m2l_pairs"], row["p2p_pairs"] = int(st["m2l_pairs"]), int(st["p2p_pairs"])
                api.check("opt", api.ps_set_option(s.ctx, api.OPT_STATS, 0))
                ts = []
                for _ in range(max(a.steps, 3)):
                    s.load(x, y, z, m)
                    api.check("sync", L.ps_hip_synchronize())
                    t0 = time.perf_counter()
                    s.tick(mmode)  # prepare + calc_forces, synchronous
                    ts.append(time.perf_counter() - t0)
                row["gpu_s"] = float(np.median(ts))
                # host arrays, as the benchmark passes them
                hx, hy, hz, hm = (np.array(t, np.float32) for t in (x, y, z, m))
                hf = [np.zeros(n, np.float32) for _ in range(3)]
                hid, hcodes = np.arange(n, dtype=np.uint32), np.zeros(n, np.uint32)
                harrs = api.make_arrs(hx, hy, hz, hm, hf[0], hf[1], hf[2], hid, n)
                th = []
                for _ in range(4):
                    hx[:], hy[:], hz[:], hid[:] = x, y, z, np.arange(n, dtype=np.uint32)
                    t0 = time.perf_counter()
                    rc, lo, _, rg = api.ps_prepare_particles(harrs, hcodes, mmode)
                    api.check("ps_prepare_particles", rc)
                    c = np.float32(np.float32(lo) + np.float32(rg) / np.float32(2.0))
                    api.check("ps_calc_forces", api.ps_calc_forces(
                        s.ctx, harrs, hcodes, c, c, c, np.float32(rg) / np.float32(2.0)))
                    th.append(time.perf_counter() - t0)
                row["gpu_host_arrays_s"] = float(np.median(th[1:]))

                def ref_fn(rr, pos):
                    with api.Solver(n) as t:
                        t.load(rr["x"], rr["y"], rr["z"], rr["m"])
                        return t.direct_f64(pos)
                if O.ref_available("shipped_mt"):
                    for thr in cpu_thr:
                        build = "shipped_mt" if thr <= 33 else "shipped_mt_wide"
                        secs = O.ref_time_ticks(x, y, z, m, theta=theta, thrd_cnt=thr, n_ticks=4,
                                                build=build)
                        row[f"ref_{thr}thr_s"] = float(np.median(secs[1:]))
                    rr = O.ref_tick(x, y, z, m, theta=theta, thrd_cnt=8, build="shipped_mt",
                                    dump=False)
                    fr = np.stack([rr["fx"], rr["fy"], rr["fz"]], 1)
                    row["ref_fp64_rms_err"] = fp64_error(n, fr, rr["id"],
                                                         lambda pos: ref_fn(rr, pos))[0]
            if theta == 1.0:
                d, p1, p8 = PUBLISHED_SWEEP[n]
                row.update(published_direct_s=d, published_1core_s=p1, published_8core_s=p8)
            row["gpu_particles_per_s"] = round(n / row["gpu_s"], 1)
            rows.append(row)
            log(f"bench: sweep theta={theta} n={n}: gpu {1e3 * row['gpu_s']:.3f} ms, host arrays "
                f"{1e3 * row['gpu_host_arrays_s']:.3f} ms, ref 1 thr "
                f"{1e3 * row.get('ref_1thr_s', float('nan')):.2f} ms, FP64 rms gpu "
                f"{row['gpu_fp64_rms_err']:.3e} ref {row.get('ref_fp64_rms_err', float('nan')):.3e}")
    head = next(r for r in rows if r["n"] == 200000 and r["theta"] == 1.0)
    line = dict(metric=METRIC, value=head["gpu_particles_per_s"], unit="particles/s", n_gpus=1,
                steps=max(a.steps, 3), warmup=1, ms_per_step=round(1e3 * head["gpu_s"], 4),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic",
                config=dict(workload="sweep", sizes=list(SWEEP_SIZES), thetas=[1.0, 2.0],
                            description="tests/benchmark.c: uniform [0,100)^3, glibc rand() "
                                        "seed 7331, mass 1, PS_MULTITHREADING traversal",
                            step="prepare + calc_forces (device-resident arrays; host arrays "
                                 "in gpu_host_arrays_s)", morton=a.morton),
                cpu_threads=cpu_thr, host=info, sweep=rows,
                note="published_* are the reference's plot values at theta=1.0 on unstated "
                     "hardware; ref_*thr_s are the reference built here, timed on this host")
    print(json.dumps(line), flush=True)
    return 0


def step_roofline(n, nodes, m2l, p2p, stages_ms, ms_step):
    """The whole step against its roofline: t_roof = sum over the stages of
    max(algorithmic bytes / 8 TB/s, algorithmic FLOP / 157.3 TFLOP/s), with
    N particles, M nodes, L M2L pairs and P P2P particle pairs; frac =
    t_roof / the measured step time.  L2P is fused into the P2P kernel, so
    its 16 B per particle are priced with P2P; the frame step (gather +
    Euler + id reset) is part of the timed step and of t_roof."""
    N, M, L, P = float(n), float(nodes), float(m2l), float(p2p)
    work = (("prepare", 16 * N, 0.0, None), ("sort", 48 * N, 0.0, "sort"),
            ("build", 4 * N + 32 * M, 0.0, "build"), ("upward", 16 * N + 32 * M, 0.0, "upward"),
            ("m2l", 0.0, FLOP_M2L * L, "traverse"), ("l2l", 24 * M, 0.0, "l2l"),
            ("p2p_l2p", 16 * N, FLOP_P2P * P, "p2p"), ("frame", 68 * N, 0.0, None))
    stages = {}
    total = 0.0
    for name, by, fl, key in work:
        roof = max(by / (HBM_PEAK_GBS * 1e9), fl / (FP32_PEAK_TFLOPS * 1e12)) * 1e3
        total += roof
        d = dict(bytes=by, flop=fl, roof_ms=round(roof, 5),
                 bound="hbm" if by / (HBM_PEAK_GBS * 1e9) >= fl / (FP32_PEAK_TFLOPS * 1e12)
                 else "valu")
        if key and stages_ms.get(key):
            d["measured_ms"] = round(stages_ms[key], 4)
            d["frac"] = round(roof / stages_ms[key], 4)
        stages[name] = d
    return dict(t_roof_ms=round(total, 5), t_step_ms=round(ms_step, 4),
                frac=round(total / ms_step, 5), stages=stages,
                peaks="8 TB/s HBM, 157.3 TFLOP/s FP32 (published MI355X)")


def measure_plain(a, name, steps, warmup, cpu_ticks, host_ticks, do_cpu, cpu_threads=None):
    """One plain workload: the timed steps, the dominant kernel's roofline,
    the whole step's roofline, the FP64 error with plain and compensated
    accumulation, the host-pointer rate and the reference CPU baseline."""
    from polesitter_amd import api
    n, x, y, z, m, v, desc = workload(name)
    log(f"bench: {api.device_name()} workload {name}: {desc}")
    L = api.hip_lib()
    s = api.Solver(n, theta=a.theta, kahan=a.kahan, stats=True)
    api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_TRAV_LANES, a.trav_lanes))
    api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_P2P_BIG, a.p2p_big))
    if a.trav_blocks:
        api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_TRAV_BLOCKS, a.trav_blocks))
    if a.trav_segments:
        api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_TRAV_SEGMENTS,
                                                     a.trav_segments))
    s.load(x, y, z, m)
    vel = [api.DeviceArray.from_numpy(vv) for vv in v]
    tmp = api.DeviceArray(3 * n, np.float32)
    dt = 1e-3
    mmode = api.MORTON_AVX2 if a.morton == "avx2" else api.MORTON_TRUNC

    def step():  # ids are 0..n-1 on entry (load, or the previous frame step)
        s.tick(mmode)
        s.frame_step(vel, tmp, dt)
        return s.stats()

    # error of the first tick (initial state), untimed
    rec = {}
    if not a.no_error:
        s.reset_ids()
        s.tick(mmode)
        r = s.result()
        f = np.stack([r["fx"], r["fy"], r["fz"]], 1)
        err = fp64_error(n, f, r["id"], s.direct_f64)
        del r, f
        log(f"bench: GPU FP64 error on 4096 targets rms={err[0]:.3e} max={err[1]:.3e}")
        # the same tick with the other accumulation (compensated vs plain FP32)
        # in the far field and P2P sums, on the same targets
        s.load(x, y, z, m)
        api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_KAHAN, 0 if a.kahan else 1))
        s.reset_ids()
        s.tick(mmode)
        r2 = s.result()
        f2 = np.stack([r2["fx"], r2["fy"], r2["fz"]], 1)
        err2 = fp64_error(n, f2, r2["id"], s.direct_f64)
        del r2, f2
        ms_other = s.stats()["ms_total"]
        api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_KAHAN, 1 if a.kahan else 0))
        s.reset_ids()
        s.tick(mmode)
        ms_this = s.stats()["ms_total"]
        this, other = ("kahan", "plain") if a.kahan else ("plain", "kahan")
        rec["fp64_rms_err"], rec["fp64_max_err"] = err
        rec["accumulation"] = {this: dict(fp64_rms_err=err[0], fp64_max_err=err[1],
                                          ms_tick=round(ms_this, 4)),
                               other: dict(fp64_rms_err=err2[0], fp64_max_err=err2[1],
                                           ms_tick=round(ms_other, 4))}
        log(f"bench: {other} accumulation: FP64 rms={err2[0]:.6e} max={err2[1]:.4e}")
        s.load(x, y, z, m)

    # warm-up steps count the P2P pairs (extra reductions); the timed steps
    # record four events each (the traversal and the P2P stage, for the
    # kernel roofline), which takes its pair counts from the last warm-up
    # step; the stage breakdown comes from a few more steps after the timed
    # ones, with every stage event recorded
    counted = None
    warmup = max(warmup, 1)
    for _ in range(warmup):
        counted = step()
    api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_STATS, 3))
    api.check("sync", L.ps_hip_synchronize())
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(step())
    api.check("sync", L.ps_hip_synchronize())
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / steps
    value = n * steps / wall
    api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_STATS, 2))
    staged = [step() for _ in range(min(steps, 5))]

    mean = lambda k: float(np.mean([st[k] for st in stats]))
    mean_staged = lambda k: float(np.mean([st[k] for st in staged]))
    kern = {
        "trav_step": dict(ms=mean("ms_trav_steps"), flop=FLOP_M2L * mean("m2l_pairs")),
        "p2p": dict(ms=mean("ms_p2p"), flop=FLOP_P2P * counted["p2p_pairs"]),
    }
    dom = max(kern, key=lambda k: kern[k]["ms"])
    ach = kern[dom]["flop"] / (kern[dom]["ms"] * 1e-3) / 1e12
    traffic = measured_traffic(name, dom)
    roof = dict(kernel=dom, bound="valu", achieved=round(ach, 4), peak=FP32_PEAK_TFLOPS,
                unit="TFLOP/s", frac=round(ach / FP32_PEAK_TFLOPS, 6),
                traffic=None if traffic is None else round(traffic),
                algorithmic=("41 FLOP per M2L pair" if dom == "trav_step"
                             else "20 FLOP per particle pair"),
                ms_per_tick=round(kern[dom]["ms"], 4),
                launch="all of the family's dispatches in one tick (achieved = FLOP per tick / "
                       "their summed duration from HIP events on the engine's stream)",
                note="FP32 VALU work (no MFMA); peak is the published FP32 figure, the "
                     "non-packed VALU ceiling is half of it; traffic = HBM-side bytes per tick "
                     "(read requests by size + WRITE_SIZE, each scaled by a streaming kernel "
                     "of known size) from profiles/traffic_" + name + ".md")
    if traffic is not None:
        roof["traffic_GBps"] = round(traffic / (kern[dom]["ms"] * 1e-3) / 1e9, 1)
    stages_ms = {k: round(mean_staged("ms_" + k), 4) for k in
                 ("sort", "build", "upward", "traverse", "l2l", "p2p", "total", "trav_steps")}
    counts = dict(nodes=int(stats[-1]["n_nodes"]), m2l_pairs=int(mean("m2l_pairs")),
                  p2p_pairs=int(counted["p2p_pairs"]), iterations=int(stats[-1]["iterations"]))
    roof["step"] = step_roofline(n, counts["nodes"], counts["m2l_pairs"], counts["p2p_pairs"],
                                 stages_ms, ms_step)
    roof["tick_frac"] = roof["step"]["frac"]
    out = dict(value=round(value, 1), ms_per_step=round(ms_step, 4), steps=steps, warmup=warmup,
               config=dict(workload=name, description=desc, n_particles=n, theta=a.theta,
                           traversal="MT depth-2 jobs (PS_MULTITHREADING)",
                           step="tick + id-gather + semi-implicit Euler dt=1e-3",
                           kahan=bool(a.kahan), trav_lanes=a.trav_lanes or "auto",
                           p2p_big=a.p2p_big, morton=a.morton, pointers="device-resident"),
               roofline=roof, stages_ms=stages_ms, counts=counts)
    out.update(rec)
    if host_ticks > 0:
        # the same tick through host arrays: ps_calc_forces stages them through
        # device buffers and copies the sorted state and forces back
        hx, hy, hz, hm = (np.ascontiguousarray(t, np.float32).copy() for t in (x, y, z, m))
        hf = [np.zeros(n, np.float32) for _ in range(3)]
        hid = np.arange(n, dtype=np.uint32)
        hcodes = np.zeros(n, np.uint32)
        harrs = api.make_arrs(hx, hy, hz, hm, hf[0], hf[1], hf[2], hid, n)
        api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_STATS, 0))

        def host_tick():
            hid[:] = np.arange(n, dtype=np.uint32)
            rc, lo, _, rg = api.ps_prepare_particles(harrs, hcodes, mmode)
            api.check("ps_prepare_particles", rc)
            c = np.float32(np.float32(lo) + np.float32(rg) / np.float32(2.0))
            api.check("ps_calc_forces", api.ps_calc_forces(s.ctx, harrs, hcodes, c, c, c,
                                                           np.float32(rg) / np.float32(2.0)))

        host_tick()
        t0 = time.perf_counter()
        for _ in range(host_ticks):
            host_tick()
        th = (time.perf_counter() - t0) / host_ticks
        out["host_pointers"] = dict(value=round(n / th, 1), unit="particles/s",
                                    ms_per_tick=round(1e3 * th, 4), ticks=host_ticks,
                                    note="prepare + calc_forces on pageable host arrays "
                                         "(uploads, tick, sorted state and forces copied "
                                         "back); not the headline value")
    s.close()
    if do_cpu:
        info = host_cpus()
        threads = cpu_threads or cpu_thread_counts(info)

        def ref_fn(r, pos):
            # FP64 direct sum on the device over the reference's sorted state
            with api.Solver(n) as t:
                t.load(r["x"], r["y"], r["z"], r["m"])
                return t.direct_f64(pos)
        try:
            cb = cpu_baseline(n, x, y, z, m, a.theta, cpu_ticks, ref_fn, threads, info)
            out["cpu_baseline"] = cb
            out["gpu_over_cpu"] = round(value / cb["value"], 2)
            if "by_threads" in cb and "1" in cb["by_threads"]:
                out["gpu_over_cpu_1thread"] = round(
                    value / cb["by_threads"]["1"]["particles_per_s"], 1)
        except Exception as exc:  # report, never hide, a failed baseline
            out["cpu_baseline"] = dict(value=None, error=str(exc))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="plummer1m")
    ap.add_argument("--theta", type=float, default=2.0)
    ap.add_argument("--kahan", action="store_true")
    ap.add_argument("--trav-lanes", type=int, default=0, choices=(0, 8, 16, 32, 64))
    ap.add_argument("--trav-blocks", type=int, default=0,
                    help="workgroups of one traversal iteration (a multiple of 8; 0: one per CU)")
    ap.add_argument("--trav-segments", type=int, default=0,
                    help="traversal work pieces per workgroup and iteration (0: the default)")
    ap.add_argument("--morton", default="avx2", choices=("avx2", "trunc"),
                    help="quantization of ps_prepare_particles: the reference's AVX2 path "
                         "(round to nearest, the shipped build the CPU baseline runs) or its "
                         "scalar path (truncation)")
    ap.add_argument("--p2p-big", type=int, default=24,
                    help="P2P wavefront-per-leaf threshold in particles (0 = off)")
    ap.add_argument("--cpu-ticks", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-16m", action="store_true",
                    help="plummer1m: skip the 16M two-Plummer record (the metric's second size)")
    ap.add_argument("--mlp-caps", default="1,8,16,32,64",
                    help="mlp4m: leaf capacities of the plain settings measured")
    ap.add_argument("--mlp-thetas", default="2.0,1.9,1.8",
                    help="mlp4m: thetas of the plain settings measured (smaller is coarser)")
    ap.add_argument("--mlp-candidates", type=int, default=3,
                    help="mlp4m: coarse settings that get a trained correction")
    ap.add_argument("--mlp-ticks", type=int, default=200, help="mlp4m: training ticks")
    ap.add_argument("--mlp-err-tol", type=float, default=0.01,
                    help="mlp4m: relative FP64-error tolerance of 'equal error'")
    ap.add_argument("--no-tune", action="store_true",
                    help="ensemble512: skip the autotuner (reference setting)")
    ap.add_argument("--host-ticks", type=int, default=3,
                    help="ticks timed with host (pageable numpy) arrays, PCIe included; 0 = skip")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {a.gpus} requested)", file=sys.stderr)
        return 2

    if a.workload == "ensemble512":
        return bench_ensemble(a)
    if a.workload == "mlp4m":
        return bench_mlp(a)
    if a.workload == "sweep":
        return bench_sweep(a)
    r = measure_plain(a, a.workload, a.steps, a.warmup, a.cpu_ticks, a.host_ticks,
                      not a.no_cpu_baseline)
    line = dict(metric=METRIC, value=r.pop("value"), unit="particles/s", n_gpus=1,
                steps=r.pop("steps"), warmup=r.pop("warmup"), ms_per_step=r.pop("ms_per_step"),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32",
                data="synthetic", config=r.pop("config"))
    line.update(r)
    if a.workload == "plummer1m" and not a.no_16m:
        # the metric's second size: N = 16,777,216 two Plummers, fewer steps
        # (each is ~10x the 1M step) and a bounded CPU sample at the thread
        # count that was fastest at 1M
        cb = line.get("cpu_baseline", {})
        thr = [cb["cores"]] if cb.get("cores") else None
        r16 = measure_plain(a, "twoplummer16m", max(3, min(a.steps, 10)), min(a.warmup, 2),
                            1, min(a.host_ticks, 1), not a.no_cpu_baseline, cpu_threads=thr)
        line["sizes"] = {"twoplummer16m": r16}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
