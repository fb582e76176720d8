"""The benchmark configurations at their real sizes on the device.

* Two Plummer spheres, N = 16,777,216 (a = 1, N/2 each at (+-5, 0, 0), bulk
  velocities (-+0.3, 0, 0)): five ticks with the caller's frame step between
  them, each checked for the sorted order, the id permutation, leaves that
  tile [0, N) and the root monopole; the tick reproduced bit for bit from the
  same state; plain and compensated (Kahan) FP32 accumulation against the
  FP64 direct sum on 4096 seeded targets (tests/accuracy.c:505-531).
* The 512-system ensemble (N_s log-uniform in [2048, 262144], 28.6M
  particles) as one batched tick recorded as a HIP graph: every system
  sorted and permuted, the replay bit-identical to the direct run, and eight
  systems (the largest, the smallest and six more) against the oracle.
* The learned far-field correction at N = 4,194,304 Plummer: trained on the
  device with a short schedule at the coarse setting, applied in the tick on
  a third realisation, where it must beat the plain coarse tick's FP64 error.

Exact oracle parity of a whole tick is checked up to N = 1,048,576
(test_gpu_scale.py); at these sizes the checks are the size-independent
properties above."""
import numpy as np
import pytest

from gpu_helpers import FORCE_RMS_TOL, forces, log, rms_rel
from polesitter_amd import api, mlp
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("codes", "id", "x", "y", "z", "m", "fx", "fy", "fz")


def fp64_rms(s, r, n, seed=4096):
    """RMS relative force error on 4096 seeded original indices, matched
    through id[] (tests/accuracy.c:505-531), FP64 direct sum on the device."""
    tg = W.sample_targets(n, 4096, seed=seed)
    inv = np.empty(n, np.int64)
    inv[r["id"]] = np.arange(n)
    pos = np.ascontiguousarray(inv[tg], np.uint32)
    ref = s.direct_f64(pos)
    f = np.stack([r["fx"], r["fy"], r["fz"]], 1)[pos].astype(np.float64)
    return rms_rel(f, ref)


def check_tick_structure(r, tree, n, total_mass):
    codes = r["codes"]
    assert np.all(codes[1:] >= codes[:-1]), "codes not sorted"
    assert np.array_equal(np.sort(r["id"]), np.arange(n, dtype=np.uint32)), "ids not a permutation"
    lf = tree["leaf"]
    first = tree["first"][lf].astype(np.int64)
    end = tree["nend"][lf].astype(np.int64)
    order = np.argsort(first)
    assert first[order][0] == 0 and end[order][-1] == n
    assert np.array_equal(first[order][1:], end[order][:-1]), "leaves do not tile [0, n)"
    assert tree["n_nodes"] > 0
    assert abs(float(tree["mp"][0, 0]) - total_mass) < 1e-4 * total_mass, tree["mp"][0]
    assert np.all(np.isfinite(r["fx"])) and np.all(np.isfinite(r["fy"])) and \
        np.all(np.isfinite(r["fz"]))


def test_twoplummer_16m_five_ticks_and_accumulation():
    n = 1 << 24
    x, y, z, m, vx, vy, vz = W.two_plummers(n, seed=1)
    total_mass = float(m.astype(np.float64).sum())
    mode = api.MORTON_AVX2
    with api.Solver(n, theta=2.0, stats=True) as s:
        # the first tick, twice from the same state: bit for bit the same
        log("16M: tick from the initial state")
        s.load(x, y, z, m)
        s.tick(mode)
        a = s.result()
        st_a = s.stats()
        e_plain = fp64_rms(s, a, n)
        log(f"16M: nodes {st_a['n_nodes']} m2l {st_a['m2l_pairs']} p2p {st_a['p2p_pairs']} "
            f"FP64 RMS {e_plain:.6e}")
        assert st_a["m2l_pairs"] > 0 and st_a["p2p_pairs"] > 0
        s.load(x, y, z, m)
        s.tick(mode)
        b = s.result()
        st_b = s.stats()
        for f in FIELDS:
            assert np.array_equal(a[f], b[f]), f
        for k in ("n_nodes", "m2l_pairs", "p2p_pairs", "p2p_leaf_pairs"):
            assert st_a[k] == st_b[k], k
        del b
        # compensated accumulation of the far-field and P2P sums
        api.check("opt", api.ps_set_option(s.ctx, api.OPT_KAHAN, 1))
        s.load(x, y, z, m)
        s.tick(mode)
        c = s.result()
        e_kahan = fp64_rms(s, c, n)
        assert np.array_equal(c["id"], a["id"])
        log(f"16M: FP64 RMS plain {e_plain:.9e} Kahan {e_kahan:.9e}")
        assert e_plain < 0.1 and e_kahan < 0.1
        # the first-order expansion's truncation dominates either way
        assert abs(e_plain - e_kahan) < 1e-6, (e_plain, e_kahan)
        del c, a
        api.check("opt", api.ps_set_option(s.ctx, api.OPT_KAHAN, 0))
        # five ticks with the frame step (id-gather, semi-implicit Euler
        # dt = 1e-3, ids reset) between them
        s.load(x, y, z, m)
        vel = [api.DeviceArray.from_numpy(v) for v in (vx, vy, vz)]
        tmp = api.DeviceArray(3 * n, np.float32)
        prev = None
        for t in range(5):
            log(f"16M: tick {t + 1}/5")
            s.tick(mode)
            r = s.result()
            st = s.stats()
            check_tick_structure(r, s.tree(), n, total_mass)
            assert st["m2l_pairs"] > 0 and st["p2p_pairs"] > 0
            if prev is not None:  # the particles moved: the tree changed a little
                assert abs(st["n_nodes"] - prev) < 0.05 * prev
            prev = st["n_nodes"]
            del r
            s.frame_step(vel, tmp, 1e-3)


def test_ensemble512_graph_tick(oracle):
    systems = W.ensemble(512, seed=1)
    sizes = np.array([len(t[0]) for t in systems])
    n = int(sizes.sum())
    mode = api.MORTON_AVX2
    log(f"ensemble: {len(systems)} systems, {n} particles")
    other = W.plummer(50_000, seed=71)
    with api.Batch(n, len(systems), theta=2.0, graph=True) as b:
        b.load(systems)
        b.tick(mode)  # runs directly, then records the graph
        st0 = b.stats()
        assert st0["used_graph"] == 0 and st0["n_systems"] == 512
        first = b.result()
        log(f"ensemble: direct tick {st0['ms_tick']:.2f} ms, nodes {st0['n_nodes']}, "
            f"m2l {st0['m2l_pairs']}")
        for k in range(len(systems)):
            g = b.system(first, k)
            assert np.all(g["codes"][1:] >= g["codes"][:-1]), k
            assert np.array_equal(np.sort(g["id"]), np.arange(sizes[k], dtype=np.uint32)), k
            assert np.all(np.isfinite(g["fx"])), k
        # replays, with unrelated device work in between
        for rep in range(2):
            with api.Solver(len(other[0])) as s2:
                s2.load(*other)
                s2.tick()
            b.restore()
            b.tick(mode)
            st = b.stats()
            assert st["used_graph"] == 1, st
            assert st["m2l_pairs"] == st0["m2l_pairs"] and st["n_nodes"] == st0["n_nodes"]
            again = b.result()
            for f in FIELDS:
                assert np.array_equal(again[f], first[f]), (rep, f)
            del again
            log(f"ensemble: replay {rep + 1} {st['ms_tick']:.2f} ms")
        # eight systems against the oracle: the largest, the smallest, six more
        rng = np.random.default_rng(512)
        pick = {int(np.argmax(sizes)), int(np.argmin(sizes))}
        pick |= set(int(k) for k in rng.choice(len(systems), 6, replace=False))
        for k in sorted(pick):
            x, y, z, m = systems[k]
            o = oracle.tick(x, y, z, m, theta=2.0, mode=mode)
            g = b.system(first, k)
            for f in ("codes", "id", "x", "y", "z", "m"):
                assert np.array_equal(g[f], o[f]), (k, f)
            err = rms_rel(forces(g), forces(o))
            log(f"ensemble: system {k} (N={sizes[k]}) force RMS vs oracle {err:.2e}")
            assert err <= FORCE_RMS_TOL, (k, err)


def test_correction_at_4m_beats_plain_coarse_tick():
    # coarse setting: theta 1.9 (a smaller theta is coarser: far <=>
    # d^2 > theta^2 (hw_t + hw_s)^2), leaf capacity 16
    n = 1 << 22
    theta, cap = 1.9, 16
    zero = [np.zeros(n, np.float32)] * 3
    x, y, z, m = W.plummer(n, seed=1)
    val = W.plummer(n, seed=2) + (zero,)
    res = mlp.build_correction(x, y, z, m, zero, train_ticks=16, val_ticks=4, samples=4096,
                               theta=theta, leaf_cap=cap, steps=600, batch=8192, lr=2e-3,
                               seed=1, val_system=val, log=log)
    log(f"4M correction: validation plain {res['val_plain']:.5e} "
        f"corrected {res['val_corrected']:.5e}")
    assert res["enabled"], res
    L = api.hip_lib()
    ex, ey, ez, em = W.plummer(n, seed=3)
    errs = {}
    for use in (False, True):
        with api.Solver(n, theta=theta) as s:
            api.check("opt", api.ps_set_option(s.ctx, api.OPT_LEAF_CAP, cap))
            if use:
                api.check("set_mlp", L.ps_hip_engine_set_mlp(s.ctx.engine,
                                                             res["params"].ctypes.data))
            s.load(ex, ey, ez, em)
            s.tick(api.MORTON_AVX2)
            r = s.result()
            errs[use] = fp64_rms(s, r, n)
    log(f"4M correction, third realisation: plain coarse {errs[False]:.5e} "
        f"corrected {errs[True]:.5e}")
    assert errs[True] < 0.95 * errs[False], errs
