"""Checks of the oracle that the mathematics fixes, independent of the
reference's numbers: the FMM with every pair near is the direct sum; the
accuracy gates of tests/accuracy.c:556-568; conservation of the monopole;
the octree as its closed-form prefix definition; the stable sort as
numpy's stable argsort."""
import numpy as np
import pytest

from polesitter_amd import workloads as W


def forces(o):
    return np.stack([o["fx"], o["fy"], o["fz"]], 1)


def ref64(oracle, o):
    return oracle.direct_f64(o["x"], o["y"], o["z"], o["m"])


@pytest.mark.parametrize("gen", [W.ref_fill_uniform, W.ref_fill_two_clusters])
def test_all_near_is_direct_sum(oracle, gen):
    x, y, z, m = gen(512)
    o = oracle.tick(x, y, z, m, theta=1e6)
    assert o["counts"]["m2l_pairs"] == 0
    rms, mx = oracle.error_metrics(forces(o), ref64(oracle, o))
    assert rms < 2e-6 and mx < 2e-5, (rms, mx)


@pytest.mark.parametrize("gen,name", [(W.ref_fill_uniform, "uniform-512"),
                                      (W.ref_fill_two_clusters, "two-clusters-512")])
def test_accuracy_gates(oracle, gen, name):
    # tests/accuracy.c:556-568: RMS relative error <= 0.10 at theta=2, MT traversal
    x, y, z, m = gen(512)
    o = oracle.tick(x, y, z, m, theta=2.0)
    rms, _ = oracle.error_metrics(forces(o), ref64(oracle, o))
    assert rms <= 0.10, (name, rms)


def test_error_falls_with_theta(oracle):
    x, y, z, m = W.plummer(4000, seed=4)
    errs = []
    for th in (1.5, 2.0, 3.0, 5.0):
        o = oracle.tick(x, y, z, m, theta=th)
        errs.append(oracle.error_metrics(forces(o), ref64(oracle, o))[0])
    assert errs == sorted(errs, reverse=True), errs  # larger theta = more accurate


def test_monopole_and_dipole_invariants(oracle):
    x, y, z, m = W.plummer(5000, seed=9)
    o = oracle.tick(x, y, z, m, theta=2.0)
    t = o["tree"]
    root = t["mp"][0]
    assert abs(root[0] - m.astype(np.float64).sum()) < 1e-4
    c = t["geom"][0, :3].astype(np.float64)
    pos = np.stack([o["x"], o["y"], o["z"]], 1).astype(np.float64)
    dip = (o["m"][:, None] * (pos - c)).sum(0)
    np.testing.assert_allclose(root[1:], dip, atol=1e-4)


def test_tree_matches_prefix_definition(oracle):
    # node set = root plus every (d, prefix_d) of a particle whose depth-(d-1)
    # run holds >= 2 particles, d <= 10; leaves: d == 10 or 1 particle
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(1, 200))
        base = rng.integers(0, 1 << 30, 4).astype(np.uint64)
        codes = np.sort((base[rng.integers(0, 4, n)] ^ rng.integers(0, 1 << rng.integers(1, 30), n).astype(np.uint64)) & ((1 << 30) - 1)).astype(np.uint32)
        t = oracle.Tree(oracle.node_cap(n))
        nn = oracle.lib().or_build(n, codes, 0.0, 0.0, 0.0, 1.0, t.st)
        v = t.view()
        expect = {(0, 0): (0, n)}
        for d in range(1, 11):
            pref = codes >> np.uint32(30 - 3 * d)
            parent_pref = codes >> np.uint32(33 - 3 * d)
            for p in np.unique(pref):
                idx = np.nonzero(pref == p)[0]
                pp = p >> 3
                par = np.nonzero(parent_pref == pp)[0]
                if (d - 1, int(pp)) in expect and len(par) >= 2:
                    expect[(d, int(p))] = (int(idx[0]), int(idx[-1]) + 1)
        assert nn == len(expect)
        got = {}
        for k in range(nn):
            d = int(v["depth"][k])
            pref = int(codes[v["first"][k]] >> np.uint32(30 - 3 * d)) if d else 0
            got[(d, pref)] = (int(v["first"][k]), int(v["end"][k]))
            leaf = v["leaf"][k]
            assert leaf == (d == 10 or v["end"][k] - v["first"][k] <= 1)
        assert got == expect


def test_sort_matches_stable_argsort(oracle):
    x, y, z, m = W.plummer(20000, seed=2)
    codes, *_ = oracle.prepare(x, y, z)
    o = oracle.calc_forces(x, y, z, m, np.arange(20000, dtype=np.uint32), codes, 0.0, 50.0, 2.0)
    assert np.array_equal(o["id"], np.argsort(codes, kind="stable").astype(np.uint32))


def test_prepare_closed_form(oracle):
    x, y, z, m = W.uniform_cube(999, seed=8, lo=-3, hi=5)
    codes, lo, hi, rg = oracle.prepare(x, y, z)
    allv = np.concatenate([x, y, z])
    assert lo == allv.min() and hi == allv.max()
    assert rg == np.float32(np.float32(hi) - np.float32(lo))
    q = lambda v: np.clip(((v - np.float32(lo)) / np.float32(rg) * np.float32(1023)).astype(np.int32), 0, 1023)
    spread = lambda v: sum(((v >> b) & 1) << (3 * b) for b in range(10))
    expect = spread(q(x)) | (spread(q(y)) << 1) | (spread(q(z)) << 2)
    assert np.array_equal(codes, expect.astype(np.uint32))
