"""Autotuner of the tick's free settings.

Three settings change the speed of a tick without changing what it computes
beyond its accuracy:

* leaf capacity (``OPT_LEAF_CAP``): the most particles of a leaf above depth
  10.  The reference uses 1; larger leaves trade M2L interactions for exact
  P2P pairs, but apply their far field as one constant over a larger cell, so
  the error can move either way -- every candidate is checked.
* P2P tile (``OPT_P2P_BIG``): leaves of at least this many particles get a
  wavefront each (64-target rounds over a shared source list, scalar
  broadcast of the sources); smaller ones a thread per particle.
* M2L schedule (``OPT_TRAV_LANES``): lanes per traversal entry, i.e. how many
  list entries of one target are evaluated together.

The objective is the smallest device time per tick on the caller's workload,
subject to the FP64 RMS force error (``tests/accuracy.c:505-531`` metric, on
seeded targets of sampled systems, FP64 direct sum on the device) staying
within ``slack`` of the error of the reference setting.  The search is a
coordinate descent: tile and schedule at the reference leaf capacity, then
the leaf capacity, then the tile again for the chosen capacity.

    from polesitter_amd import autotune, workloads as W
    best = autotune.tune_batch(W.ensemble(512))
    best["config"]      # {"leaf_cap": ..., "p2p_big": ..., "trav_lanes": ...}
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import api
from . import workloads as W

LEAF_CAPS = (1, 2, 4, 8, 16)
P2P_BIGS = (16, 24, 32, 64)
LANES = (8, 16)


def _apply(b: "api.Batch", cfg: dict):
    b.set_option(api.OPT_LEAF_CAP, cfg["leaf_cap"])
    b.set_option(api.OPT_P2P_BIG, cfg["p2p_big"])
    b.set_option(api.OPT_TRAV_LANES, cfg["trav_lanes"])


def _batch_error(b: "api.Batch", res: dict, systems_idx, targets_per_system: int) -> float:
    """FP64 RMS error over seeded targets of the given systems (sorted state)."""
    L = api.hip_lib()
    num = den = 0.0
    for k in systems_idx:
        a, e = int(b.offsets[k]), int(b.offsets[k + 1])
        n = e - a
        g = b.system(res, k)
        tg = W.sample_targets(n, targets_per_system, seed=4096 + k)
        inv = np.empty(n, np.int64)
        inv[g["id"]] = np.arange(n)
        pos = np.ascontiguousarray(inv[tg], np.uint32)
        out = np.zeros(3 * pos.size, np.float64)
        api.check("ps_hip_direct_f64", L.ps_hip_direct_f64(
            C.c_void_p(b.x.ptr + 4 * a), C.c_void_p(b.y.ptr + 4 * a), C.c_void_p(b.z.ptr + 4 * a),
            C.c_void_p(b.m.ptr + 4 * a), n, pos.ctypes.data, pos.size, out.ctypes.data))
        ref = out.reshape(-1, 3)
        f = np.stack([g["fx"], g["fy"], g["fz"]], 1)[pos].astype(np.float64)
        num += float(((f - ref) ** 2).sum())
        den += float((ref ** 2).sum())
    return float(np.sqrt(num / den)) if den > 0 else 0.0


def _measure(b: "api.Batch", cfg: dict, mode: int, reps: int, err_sys, tps: int):
    _apply(b, cfg)
    b.restore()
    b.tick(mode)  # direct run + recording
    res = b.result()
    err = _batch_error(b, res, err_sys, tps)
    times = []
    for _ in range(reps):
        b.restore()
        b.tick(mode)
        times.append(b.stats()["ms_tick"])
    return float(np.median(times)), err


def tune_batch(systems, theta: float = 2.0, mode: int = api.MORTON_AVX2, slack: float = 0.0,
               reps: int = 3, error_systems: int = 16, targets_per_system: int = 256,
               log=None) -> dict:
    """Tune the settings for a batch of systems (list of (x, y, z, m)).

    Returns {"config", "ms_tick", "fp64_rms_err", "reference": {...},
    "table": [(config, ms, err), ...]}.  The reference setting is leaf
    capacity 1, P2P tile 24, 8 lanes per entry."""
    n = sum(len(s[0]) for s in systems)
    nsys = len(systems)
    err_sys = list(range(0, nsys, max(1, nsys // error_systems)))[:error_systems]
    table = []
    seen = {}
    with api.Batch(n, nsys, theta=theta, graph=True) as b:
        b.load(systems)

        def run(cfg):
            key = (cfg["leaf_cap"], cfg["p2p_big"], cfg["trav_lanes"])
            if key not in seen:
                ms, err = _measure(b, cfg, mode, reps, err_sys, targets_per_system)
                seen[key] = (ms, err)
                table.append((dict(cfg), ms, err))
                if log:
                    log(f"autotune: {cfg} -> {ms:.3f} ms, FP64 RMS {err:.4e}")
            return seen[key]

        ref_cfg = dict(leaf_cap=1, p2p_big=24, trav_lanes=8)
        ref_ms, ref_err = run(ref_cfg)
        limit = ref_err * (1.0 + slack)

        def better(cfg, best):
            ms, err = run(cfg)
            return err <= limit and ms < run(best)[0]

        best = dict(ref_cfg)
        for lanes in LANES:  # schedule and tile at the reference leaf capacity
            for big in P2P_BIGS:
                cfg = dict(best, trav_lanes=lanes, p2p_big=big)
                if better(cfg, best):
                    best = cfg
        for cap in LEAF_CAPS:  # leaf capacity
            cfg = dict(best, leaf_cap=cap)
            if better(cfg, best):
                best = cfg
        for big in P2P_BIGS:  # tile again for the chosen capacity
            cfg = dict(best, p2p_big=big)
            if better(cfg, best):
                best = cfg
        ms, err = run(best)
    return dict(config=best, ms_tick=ms, fp64_rms_err=err,
                reference=dict(config=ref_cfg, ms_tick=ref_ms, fp64_rms_err=ref_err),
                table=table, n_particles=n, n_systems=nsys)
