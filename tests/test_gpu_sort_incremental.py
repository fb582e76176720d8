"""The incremental sort (PS_HIP_OPT_SORT_INCREMENTAL): after a tick of the
same size, only the keys that differ from the last tick's sorted keys are
sorted and merged with the rest.  Bar: the permutation is the stable one
(numpy's stable argsort of the prepared codes) bit for bit -- including
ties between unchanged and changed keys -- the payload follows it, the whole
tick equals a tick with the full radix sort bit for bit, and more than 1/8
changed keys fall back to the full sort."""
import numpy as np
import pytest

from gpu_helpers import log
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("codes", "id", "x", "y", "z", "m", "fx", "fy", "fz")


def tick(s, mode=api.MORTON_AVX2):
    """prepare, keep the unsorted codes, calc_forces; returns (prep codes, result, stats)"""
    rc, lo, hi, rg = api.ps_prepare_particles(s.arrs, s.codes, mode)
    api.check("prepare", rc)
    prep = s.codes.numpy()
    c = np.float32(np.float32(lo) + np.float32(rg) * np.float32(0.5))
    api.check("calc", api.ps_calc_forces(s.ctx, s.arrs, s.codes, c, c, c,
                                         np.float32(rg) * np.float32(0.5)))
    return prep, s.result(), s.stats()


def check_stable(prep, before, r):
    perm = np.argsort(prep, kind="stable")
    assert np.array_equal(r["codes"], prep[perm])
    assert np.array_equal(r["id"], before["id"][perm])
    for k in ("x", "y", "z", "m"):
        assert np.array_equal(r[k], before[k][perm]), k


def state(s):
    return {k: a.numpy() for k, a in (("x", s.x), ("y", s.y), ("z", s.z), ("m", s.m),
                                       ("id", s.id))}


@pytest.mark.parametrize("frac,incremental", [(0.0005, 1), (0.05, 1), (0.11, 1), (0.3, 0)])
def test_incremental_sort_is_the_stable_order(frac, incremental):
    n = 300_001
    x, y, z, m = W.plummer(n, seed=401)
    rng = np.random.default_rng(402)
    with api.Solver(n, theta=2.0, stats=True) as s, \
            api.Solver(n, theta=2.0, stats=True) as f:
        api.check("opt", api.ps_set_option(f.ctx, api.OPT_SORT_INCREMENTAL, 0))
        for t in (s, f):
            t.load(x, y, z, m)
            tick(t)
        # move a fraction of the (now sorted) particles to new places, keep
        # the rest; ids 0..n-1 as the frame step leaves them
        st0 = state(s)
        # the particles that hold the cube's min and max stay, so the cube
        # (and the scale of every code) does too
        lo = np.minimum(np.minimum(st0["x"], st0["y"]), st0["z"])
        hi = np.maximum(np.maximum(st0["x"], st0["y"]), st0["z"])
        keep = {int(np.argmin(lo)), int(np.argmax(hi))}
        moved = np.array([k for k in rng.choice(n, int(frac * n) + 2, replace=False)
                          if int(k) not in keep][:int(frac * n)])
        for k in ("x", "y", "z"):
            st0[k][moved] = rng.uniform(-3, 3, moved.size).astype(np.float32)
        st0["id"] = np.arange(n, dtype=np.uint32)
        for t in (s, f):
            t.load(st0["x"], st0["y"], st0["z"], st0["m"], st0["id"])
        prep, r, st = tick(s)
        prep_f, rf, stf = tick(f)
    log(f"moved {moved.size}: changed keys {st['sort_changed']}, incremental "
        f"{st['sort_incremental']}")
    assert st["sort_incremental"] == incremental and stf["sort_incremental"] == 0
    assert np.array_equal(prep, prep_f)
    check_stable(prep, st0, r)
    for k in FIELDS:  # the whole tick: bit for bit the full sort's
        assert np.array_equal(r[k], rf[k]), k


def test_incremental_sort_breaks_ties_by_position():
    # particles on a coarse lattice share codes; moving some of them to
    # other lattice sites makes changed and unchanged keys tie
    n = 120_000
    rng = np.random.default_rng(403)
    sites = rng.integers(0, 40, (n, 3)).astype(np.float32) * np.float32(0.25)
    x, y, z = (np.ascontiguousarray(sites[:, k]) for k in range(3))
    m = np.full(n, 1.0 / n, np.float32)
    with api.Solver(n, theta=2.0, stats=True) as s:
        s.load(x, y, z, m)
        tick(s)
        st0 = state(s)
        moved = rng.choice(n, n // 30, replace=False)
        new = rng.integers(0, 40, (moved.size, 3)).astype(np.float32) * np.float32(0.25)
        for j, k in enumerate(("x", "y", "z")):
            st0[k][moved] = new[:, j]
        # keep the cube: the corners stay occupied so codes keep their scale
        st0["id"] = np.arange(n, dtype=np.uint32)
        s.load(st0["x"], st0["y"], st0["z"], st0["m"], st0["id"])
        prep, r, st = tick(s)
    assert st["sort_incremental"] == 1, st["sort_changed"]
    assert len(np.unique(prep)) < n // 2  # many ties
    check_stable(prep, st0, r)


def test_frame_loop_sorts_incrementally_and_equals_full_sort():
    # the caller's frame loop: tick, id-gather + Euler + id reset, five times;
    # after the first tick the sort is incremental and every tick equals the
    # full sort's bit for bit
    n = 1 << 20
    x, y, z, m, vx, vy, vz = W.two_plummers(n, seed=404)
    out = {}
    for inc in (1, 0):
        with api.Solver(n, theta=2.0, stats=True) as s:
            api.check("opt", api.ps_set_option(s.ctx, api.OPT_SORT_INCREMENTAL, inc))
            s.load(x, y, z, m)
            vel = [api.DeviceArray.from_numpy(v) for v in (vx, vy, vz)]
            tmp = api.DeviceArray(3 * n, np.float32)
            rows = []
            for t in range(5):
                before = state(s)
                prep, r, st = tick(s)
                check_stable(prep, before, r)
                rows.append((r, st["sort_incremental"], st["sort_changed"]))
                s.frame_step(vel, tmp, 1e-3)
            out[inc] = rows
    log("changed keys per tick: " + ", ".join(str(c) for _, _, c in out[1]))
    assert [i for _, i, _ in out[1]] == [0, 1, 1, 1, 1]
    for t in range(5):
        for k in FIELDS:
            assert np.array_equal(out[1][t][0][k], out[0][t][0][k]), (t, k)


def test_size_change_sorts_in_full():
    x, y, z, m = W.plummer(50_000, seed=405)
    with api.Solver(50_000, theta=2.0, stats=True) as s:
        s.load(x, y, z, m)
        tick(s)
        # the same arrays, fewer particles: no kept keys of that size
        s.n = 40_000
        s.arrs = api.make_arrs(s.x, s.y, s.z, s.m, s.fx, s.fy, s.fz, s.id, 40_000)
        before = {k: v[:40_000] for k, v in state(s).items()}
        before["id"] = np.arange(40_000, dtype=np.uint32)
        s.id.upload(np.concatenate([before["id"], np.zeros(10_000, np.uint32)]))
        prep, r, st = tick(s)
        assert st["sort_incremental"] == 0
        r = {k: v[:40_000] for k, v in r.items()}
        check_stable(prep[:40_000], before, r)
