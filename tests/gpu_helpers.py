"""Helpers of the GPU tests: run one public-API tick on the device and put
its outputs in the oracle's shape, and the parity assertions (bit-exact for
integer and index work and for values the reference computes in a fixed
order; within stated tolerances where only the summation order differs)."""
from __future__ import annotations

import sys

import numpy as np

from polesitter_amd import api

# Stated tolerances: forces and locals within 1e-5 RMS relative to the oracle
# (the GPU sums M2L and P2P terms in another order than the reference's
# depth-first recursion; everything discrete is compared bit for bit).
FORCE_RMS_TOL = 1e-5
LOCAL_RMS_TOL = 1e-5


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def gpu_tick(x, y, z, m, theta=2.0, trav=api.TRAV_MT, mode=api.MORTON_TRUNC, host=False,
             kahan=False, ids=None, tree=True, lanes=None, p2p_big=None, leaf_cap=None,
             arena_words=None, trav_blocks=None, trav_segments=None):
    """ps_prepare_particles + ps_calc_forces on the device, root = min +
    range/2 (tests/accuracy.c:493-503).  host=True passes numpy (host)
    arrays, otherwise device arrays."""
    n = len(x)
    f32 = np.float32
    x, y, z, m = (np.ascontiguousarray(a, f32).copy() for a in (x, y, z, m))
    ids = np.arange(n, dtype=np.uint32) if ids is None else np.asarray(ids, np.uint32).copy()
    with api.Solver(n, theta=theta, traversal=trav, kahan=kahan, stats=True) as s:
        if lanes is not None:
            api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_TRAV_LANES, lanes))
        if p2p_big is not None:
            api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_P2P_BIG, p2p_big))
        if leaf_cap is not None:
            api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_LEAF_CAP, leaf_cap))
        if arena_words is not None:
            api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_ARENA_WORDS,
                                                         arena_words))
        if trav_blocks is not None:
            api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_TRAV_BLOCKS,
                                                         trav_blocks))
        if trav_segments is not None:
            api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_TRAV_SEGMENTS,
                                                         trav_segments))
        if host:
            fx, fy, fz = (np.zeros(n, f32) for _ in range(3))
            codes = np.zeros(n, np.uint32)
            arrs = api.make_arrs(x, y, z, m, fx, fy, fz, ids, n)
        else:
            s.load(x, y, z, m, ids)
            arrs = s.arrs
            codes = s.codes
        rc, lo, hi, rg = api.ps_prepare_particles(arrs, codes, mode)
        api.check("ps_prepare_particles", rc)
        prep = codes.copy() if host else codes.numpy()
        rc_ = np.float32(np.float32(lo) + np.float32(rg) * np.float32(0.5))
        hw = np.float32(rg) * np.float32(0.5)
        api.check("ps_calc_forces", api.ps_calc_forces(s.ctx, arrs, codes, rc_, rc_, rc_, hw))
        if host:
            out = dict(x=x, y=y, z=z, m=m, id=ids, codes=codes, fx=fx, fy=fy, fz=fz)
        else:
            out = s.result()
        out.update(prep_codes=prep, min=lo, max=hi, range=rg, stats=s.stats())
        if tree:
            out["tree"] = s.tree()
    return out


def forces(d):
    return np.stack([d["fx"], d["fy"], d["fz"]], 1).astype(np.float64)


def rms_rel(a, b):
    den = (b * b).sum()
    return float(np.sqrt(((a - b) ** 2).sum() / den)) if den > 0 else float(np.abs(a - b).max())


def assert_sorted_state_equal(g, o):
    for k in ("codes", "id", "x", "y", "z", "m"):
        assert np.array_equal(g[k], o[k]), f"{k} differs"


def assert_tree_equal(gt, ot_first, ot_cnt, ot_leaf, ot_parent, ot_geom, ot_mp):
    """gt = exported device tree; ot_* = oracle/reference canonical arrays."""
    assert gt["n_nodes"] == len(ot_leaf), (gt["n_nodes"], len(ot_leaf))
    assert np.array_equal(gt["leaf"], ot_leaf), "leaf flags differ"
    lf = gt["leaf"]
    assert np.array_equal(gt["first"][lf], ot_first), "leaf first differs"
    assert np.array_equal((gt["nend"] - gt["first"])[lf], ot_cnt), "leaf count differs"
    par = gt["parent"].astype(np.int64)
    opar = np.asarray(ot_parent).astype(np.int64)
    assert np.array_equal(par[1:], opar[1:]), "parent differs"
    assert np.array_equal(gt["geom"], ot_geom), "centres / half widths differ"
    assert np.array_equal(gt["mp"], ot_mp), "multipoles differ"


def assert_tick_parity(g, o, counts=True):
    """g: gpu_tick output; o: oracle.tick output (same inputs)."""
    assert np.array_equal(g["prep_codes"], o["prep_codes"]), "prepare codes differ"
    assert (np.float32(g["min"]), np.float32(g["max"]), np.float32(g["range"])) == \
        (np.float32(o["min"]), np.float32(o["max"]), np.float32(o["range"]))
    assert_sorted_state_equal(g, o)
    t = o["tree"]
    lf = t["leaf"]
    assert_tree_equal(g["tree"], t["first"][lf], (t["end"] - t["first"])[lf], lf, t["parent"],
                      t["geom"], t["mp"])
    loc_err = rms_rel(g["tree"]["loc"][:, 1:].astype(np.float64), t["loc"][:, 1:].astype(np.float64))
    assert loc_err <= LOCAL_RMS_TOL, loc_err
    f_err = rms_rel(forces(g), forces(o))
    assert f_err <= FORCE_RMS_TOL, f_err
    if counts:
        st, c = g["stats"], o["counts"]
        assert st["m2l_pairs"] == c["m2l_pairs"], (st["m2l_pairs"], c["m2l_pairs"])
        assert st["p2p_leaf_pairs"] == c["p2p_leaf_pairs"]
        assert st["p2p_pairs"] == c["p2p_pairs"]
    return f_err, loc_err
