"""Autotuner of the batched tick's free settings.

Three settings change the speed of a tick without changing what it computes
beyond its accuracy:

* leaf capacity: the most particles of a leaf above depth 10.  The reference
  uses 1; larger leaves trade M2L interactions for exact P2P pairs but apply
  their far field as one constant over a larger cell, so the error can move
  either way.  It is chosen per class of system size
  (``ps_hip_batch_set_leaf_caps``): small and large systems do not answer
  the same way.
* P2P tile (``OPT_P2P_BIG``): leaves of at least this many particles get a
  wavefront each (64-target rounds over a shared source list, scalar
  broadcast of the sources); smaller ones a thread per particle.
* M2L schedule (``OPT_TRAV_LANES``): lanes per traversal entry, i.e. how many
  list entries of one target are evaluated together.  Tile and schedule are
  launch shapes of the whole forest, so they are chosen per batch.

Objective: the smallest device time per batched tick (graph replay) on the
caller's systems, subject to every size class's FP64 RMS force error
(``tests/accuracy.c:505-531`` metric, seeded targets of sampled systems of
the class, FP64 direct sum on the device) staying within ``slack`` of that
class's error under the reference setting (rounding-level differences,
below ``tol`` relative, are not accuracy changes: the lane count alone
reorders M2L sums).  The search is a coordinate descent: schedule and tile
at the reference capacities, then each class's capacity (largest share of
particles first), then the tile again.

    from polesitter_amd import autotune, workloads as W
    best = autotune.tune_batch(W.ensemble(512))
    best["config"]       # {"caps": (...per class), "p2p_big": ..., "trav_lanes": ...}
    best["system_caps"]  # the leaf capacity of every system
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import api
from . import workloads as W

LEAF_CAPS = (2, 4, 8, 16)
P2P_BIGS = (16, 24, 32, 64)
LANES = (8, 16, 32)
CLASS_EDGES = (8192, 32768, 131072)  # system sizes: [..8k), [8k, 32k), [32k, 128k), [128k..)


def size_classes(sizes, edges=CLASS_EDGES) -> np.ndarray:
    return np.searchsorted(np.asarray(edges), np.asarray(sizes), side="right")


def _apply(b: "api.Batch", cfg: dict, classes: np.ndarray):
    b.set_option(api.OPT_P2P_BIG, cfg["p2p_big"])
    b.set_option(api.OPT_TRAV_LANES, cfg["trav_lanes"])
    b.set_leaf_caps(np.asarray(cfg["caps"], np.uint8)[classes])


def _class_errors(b: "api.Batch", res: dict, err_sys, classes, targets_per_system: int):
    """FP64 RMS error per size class over seeded targets of err_sys."""
    L = api.hip_lib()
    num, den = {}, {}
    for k in err_sys:
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
        c = int(classes[k])
        num[c] = num.get(c, 0.0) + float(((f - ref) ** 2).sum())
        den[c] = den.get(c, 0.0) + float((ref ** 2).sum())
    per = {c: float(np.sqrt(num[c] / den[c])) if den[c] > 0 else 0.0 for c in num}
    total = float(np.sqrt(sum(num.values()) / sum(den.values()))) if sum(den.values()) else 0.0
    return per, total


def _measure(b, cfg, classes, mode, reps, err_sys, tps):
    _apply(b, cfg, classes)
    b.restore()
    b.tick(mode)  # direct run + recording
    res = b.result()
    per, total = _class_errors(b, res, err_sys, classes, tps)
    times = []
    for _ in range(reps):
        b.restore()
        b.tick(mode)
        times.append(b.stats()["ms_tick"])
    return float(np.median(times)), per, total


def tune_batch(systems, theta: float = 2.0, mode: int = api.MORTON_AVX2, slack: float = 0.0,
               tol: float = 1e-6, reps: int = 3, error_systems_per_class: int = 6,
               targets_per_system: int = 256, class_edges=CLASS_EDGES, log=None) -> dict:
    """Tune the settings for a batch of systems (list of (x, y, z, m)).

    Returns {"config": {"caps", "p2p_big", "trav_lanes"}, "system_caps",
    "ms_tick", "fp64_rms_err", "class_errors", "reference": {...},
    "table": [(config, ms, class errors, error), ...], "classes": edges and
    counts}.  The reference setting is leaf capacity 1 everywhere, P2P
    tile 24, 8 lanes per entry."""
    sizes = np.array([len(s[0]) for s in systems])
    classes = size_classes(sizes, class_edges)
    n_cls = len(class_edges) + 1
    present = [c for c in range(n_cls) if np.any(classes == c)]
    err_sys = []
    for c in present:  # a spread of each class's systems
        idx = np.flatnonzero(classes == c)
        pick = idx[np.linspace(0, len(idx) - 1, min(error_systems_per_class, len(idx))).astype(int)]
        err_sys += sorted(set(int(k) for k in pick))
    n = int(sizes.sum())
    table = []
    seen = {}
    with api.Batch(n, len(systems), theta=theta, graph=True) as b:
        b.load(systems)

        def key(cfg):
            return (tuple(cfg["caps"]), cfg["p2p_big"], cfg["trav_lanes"])

        def run(cfg):
            k = key(cfg)
            if k not in seen:
                ms, per, total = _measure(b, cfg, classes, mode, reps, err_sys,
                                          targets_per_system)
                seen[k] = (ms, per, total)
                table.append((dict(cfg, caps=list(cfg["caps"])), ms, per, total))
                if log:
                    log(f"autotune: caps {cfg['caps']} tile {cfg['p2p_big']} lanes "
                        f"{cfg['trav_lanes']} -> {ms:.3f} ms, FP64 RMS {total:.5e} "
                        f"(by class {', '.join(f'{c}: {per[c]:.4e}' for c in sorted(per))})")
            return seen[k]

        ref_cfg = dict(caps=(1,) * n_cls, p2p_big=24, trav_lanes=8)
        ref_ms, ref_per, ref_total = run(ref_cfg)
        limit = {c: ref_per[c] * (1.0 + slack) * (1.0 + tol) for c in ref_per}

        def ok(per):
            return all(per[c] <= limit[c] for c in per)

        def better(cfg, best):
            ms, per, _ = run(cfg)
            return ok(per) and ms < run(best)[0]

        best = dict(ref_cfg)
        for lanes in LANES:  # schedule and tile at the reference capacities
            for big in P2P_BIGS:
                cfg = dict(best, trav_lanes=lanes, p2p_big=big)
                if better(cfg, best):
                    best = cfg
        share = {c: int(sizes[classes == c].sum()) for c in present}
        for c in sorted(present, key=lambda c: -share[c]):  # capacity per class
            for cap in LEAF_CAPS:
                caps = list(best["caps"])
                caps[c] = cap
                cfg = dict(best, caps=tuple(caps))
                if better(cfg, best):
                    best = cfg
        for big in P2P_BIGS:  # tile again for the chosen capacities
            cfg = dict(best, p2p_big=big)
            if better(cfg, best):
                best = cfg
        ms, per, total = run(best)
    return dict(config=dict(best, caps=list(best["caps"])),
                system_caps=np.asarray(best["caps"], np.uint8)[classes].tolist(),
                ms_tick=ms, fp64_rms_err=total, class_errors=per,
                reference=dict(config=dict(ref_cfg, caps=list(ref_cfg["caps"])), ms_tick=ref_ms,
                               fp64_rms_err=ref_total, class_errors=ref_per),
                table=table, n_particles=n, n_systems=len(systems),
                classes=dict(edges=list(class_edges),
                             systems=[int((classes == c).sum()) for c in range(n_cls)],
                             particles=[int(sizes[classes == c].sum()) for c in range(n_cls)]))
