"""Batched ticks (ps_hip_batch_*): many independent systems in one set of
device arrays, run as one forest and replayed from a recorded HIP graph.

Bar: every system's sorted state (codes, ids, positions, masses) and forces
are bit-identical to a single-system tick of the same particles at the same
lanes per traversal entry (the same kernels visit the same lists in the same
order; the lane count sets the order in which an entry's M2L terms are
summed, and the single-system tick otherwise picks it per iteration from its
entry counts), the forest's M2L pair count
is the sum of the oracle's per-system counts, and every graph replay
reproduces the directly run tick bit for bit -- with other device work
between the replays."""
import numpy as np
import pytest

from gpu_helpers import assert_tick_parity, gpu_tick, log
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("codes", "id", "x", "y", "z", "m", "fx", "fy", "fz")
LANES = 8  # lanes per traversal entry on both sides of a comparison


def ragged_systems(seed=61):
    # sizes from 1 (a root that is a leaf) to 20k, uniform and Plummer, one
    # system of coincident particles (a single depth-10 leaf)
    sizes = [1, 2, 7, 64, 1000, 5000, 20000, 3, 333, 4096]
    out = []
    for k, n in enumerate(sizes):
        if k == 7:
            x = np.full(n, 0.25, np.float32)
            out.append((x, x.copy(), x.copy(), np.full(n, 1.0 / n, np.float32)))
        elif k % 2 == 0:
            out.append(W.uniform_cube(n, seed * 1000 + k, mass=1.0 / n))
        else:
            out.append(W.plummer(n, seed * 1000 + k, 0.5 + 0.2 * k, mass=1.0 / n))
    return out


def single(sys_, mode, trav):
    x, y, z, m = sys_
    return gpu_tick(x, y, z, m, theta=2.0, trav=trav, mode=mode, tree=False, lanes=LANES)


@pytest.mark.parametrize("mode", [api.MORTON_TRUNC, api.MORTON_AVX2])
@pytest.mark.parametrize("trav", [api.TRAV_MT, api.TRAV_ST])
def test_batch_equals_single_ticks(mode, trav):
    systems = ragged_systems()
    n = sum(len(s[0]) for s in systems)
    with api.Batch(n, len(systems), theta=2.0, traversal=trav, graph=False) as b:
        b.set_option(api.OPT_TRAV_LANES, LANES)
        b.load(systems)
        b.tick(mode)
        res = b.result()
        st = b.stats()
    log(f"batch mode={mode} trav={trav}: {st}")
    for k, sys_ in enumerate(systems):
        g = b.system(res, k)
        o = single(sys_, mode, trav)
        for f in FIELDS:
            assert np.array_equal(g[f], o[f]), (k, f)


def test_batch_leaf_cap_equals_single_ticks():
    systems = ragged_systems(seed=65)
    n = sum(len(s[0]) for s in systems)
    with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
        b.set_option(api.OPT_LEAF_CAP, 8)
        b.set_option(api.OPT_TRAV_LANES, LANES)
        b.load(systems)
        b.tick()
        res = b.result()
    for k, (x, y, z, m) in enumerate(systems):
        o = gpu_tick(x, y, z, m, theta=2.0, tree=False, leaf_cap=8, lanes=LANES)
        g = b.system(res, k)
        for f in FIELDS:
            assert np.array_equal(g[f], o[f]), (k, f)


def test_batch_pair_counts_match_oracle(oracle):
    systems = ragged_systems(seed=62)
    n = sum(len(s[0]) for s in systems)
    with api.Batch(n, len(systems), theta=2.0, graph=False) as b:
        b.load(systems)
        b.tick()
        st = b.stats()
        res = b.result()
    m2l = 0
    for k, (x, y, z, m) in enumerate(systems):
        o = oracle.tick(x, y, z, m, theta=2.0)
        m2l += o["counts"]["m2l_pairs"]
        g = b.system(res, k)
        # codes, ids and payload bit-exact; forces within the parity tolerance
        for f in ("codes", "id", "x", "y", "z", "m"):
            assert np.array_equal(g[f], o[f]), (k, f)
        den = (np.stack([o["fx"], o["fy"], o["fz"]], 1).astype(np.float64) ** 2).sum()
        num = ((np.stack([g["fx"], g["fy"], g["fz"]], 1).astype(np.float64)
                - np.stack([o["fx"], o["fy"], o["fz"]], 1)) ** 2).sum()
        rel = np.sqrt(num / den) if den > 0 else np.sqrt(num)
        assert rel <= 1e-5, (k, rel)
    assert st["m2l_pairs"] == m2l, (st["m2l_pairs"], m2l)


def test_graph_replays_match_direct_run():
    systems = W.ensemble(24, seed=3, lo=512, hi=20000)
    n = sum(len(s[0]) for s in systems)
    other = W.plummer(30_000, seed=63)
    with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
        b.load(systems)
        b.tick()  # runs directly and records the graph
        first = b.result()
        assert b.stats()["used_graph"] == 0
        for r in range(3):
            gpu_tick(*other, tree=False)  # unrelated device work between replays
            b.restore()
            b.tick()
            st = b.stats()
            assert st["used_graph"] == 1, st
            again = b.result()
            for f in FIELDS:
                assert np.array_equal(again[f], first[f]), (r, f)
        log(f"replays: {b.stats()}")


def test_graph_rerecorded_after_layout_change():
    a = ragged_systems(seed=64)
    with api.Batch(40_000, 16, theta=2.0, graph=True) as b:
        b.load(a)
        b.tick()
        b.restore()
        b.tick()
        assert b.stats()["used_graph"] == 1
        b.load(a[:5])  # other offsets: a new recording
        b.tick()
        assert b.stats()["used_graph"] == 0
        r1 = b.result()
        b.restore()
        b.tick()
        assert b.stats()["used_graph"] == 1
        r2 = b.result()
        for f in FIELDS:
            assert np.array_equal(r1[f], r2[f]), f


def test_batch_rejects_bad_layouts():
    with api.Batch(1000, 4) as b:
        o = np.array([0, 10, 10, 20], np.uint32)  # an empty system
        L = api.hip_lib()
        rc = L.ps_hip_batch_tick(b.h, b.x.ptr, b.y.ptr, b.z.ptr, b.m.ptr, b.fx.ptr, b.fy.ptr,
                                 b.fz.ptr, b.id.ptr, b.codes.ptr, o.ctypes.data, 3, 0)
        assert rc == api.PS_EINVAL
        o = np.array([0, 10, 2000], np.uint32)  # more than max_particles
        rc = L.ps_hip_batch_tick(b.h, b.x.ptr, b.y.ptr, b.z.ptr, b.m.ptr, b.fx.ptr, b.fy.ptr,
                                 b.fz.ptr, b.id.ptr, b.codes.ptr, o.ctypes.data, 2, 0)
        assert rc == api.PS_EINVAL


@pytest.mark.parametrize("mode", [api.MORTON_TRUNC, api.MORTON_AVX2])
def test_batch_arena_overflow_equals_single_ticks(mode):
    # a list arena of one word per particle overflows in the first tick: the
    # arenas grow and the tick runs again from the traversal over the sorted
    # arrays and the forest it already has.  Sizes that are not multiples of
    # 8 make the AVX2 quantisation depend on each particle's position in its
    # system, so running the prepare and the sort a second time would change
    # codes; the result must equal single ticks bit for bit, and so must the
    # replay of the graph recorded after the overflow.
    sizes = [333, 5001, 20003, 7, 4099, 64]
    systems = [W.plummer(k, seed=90 + j, mass=1.0 / k) if j % 2 else
               W.uniform_cube(k, seed=90 + j, mass=1.0 / k) for j, k in enumerate(sizes)]
    n = sum(sizes)
    with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
        b.set_option(api.OPT_TRAV_LANES, LANES)
        b.set_option(api.OPT_ARENA_WORDS, 1)
        b.load(systems)
        b.tick(mode)
        res = b.result()
        b.restore()
        b.tick(mode)
        assert b.stats()["used_graph"] == 1
        rep = b.result()
    for f in FIELDS:
        assert np.array_equal(rep[f], res[f]), f
    for k, sys_ in enumerate(systems):
        g = b.system(res, k)
        o = single(sys_, mode, api.TRAV_MT)
        for f in FIELDS:
            assert np.array_equal(g[f], o[f]), (k, f)


@pytest.mark.parametrize("mode", [api.MORTON_TRUNC, api.MORTON_AVX2])
def test_batch_per_system_leaf_caps_equal_single_ticks(mode):
    # ps_hip_batch_set_leaf_caps: system k builds the tree of a single tick
    # with PS_HIP_OPT_LEAF_CAP = caps[k], and everything after it follows
    systems = ragged_systems(seed=67)
    caps = [1, 2, 4, 8, 16, 3, 64, 1, 5, 12]
    n = sum(len(s[0]) for s in systems)
    with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
        b.set_option(api.OPT_TRAV_LANES, LANES)
        b.set_leaf_caps(caps)
        b.load(systems)
        b.tick(mode)
        res = b.result()
        b.restore()
        b.tick(mode)  # the graph replay keeps the capacities
        assert b.stats()["used_graph"] == 1
        rep = b.result()
        # fewer capacities than systems is refused
        b.set_leaf_caps(caps[:3])
        with pytest.raises(api.PsError):
            b.tick(mode)
    for f in FIELDS:
        assert np.array_equal(rep[f], res[f]), f
    for k, (x, y, z, m) in enumerate(systems):
        o = gpu_tick(x, y, z, m, theta=2.0, mode=mode, tree=False, leaf_cap=caps[k],
                     lanes=LANES)
        g = b.system(res, k)
        for f in FIELDS:
            assert np.array_equal(g[f], o[f]), (k, caps[k], f)
