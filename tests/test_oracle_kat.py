"""The reference's own known-answer tests (tests/main.c, tests/accuracy.c)
restated against the oracle's operators."""
import math

import numpy as np
import pytest

THETA_KAT = 3.4641  # sqrt(12), tests/main.c:11


def test_morton_encoding(oracle):
    # tests/main.c:50-72
    assert oracle.morton_encode(1, 0, 0) == 1
    assert oracle.morton_encode(0, 1, 0) == 2
    assert oracle.morton_encode(0, 0, 1) == 4
    assert oracle.morton_encode(3, 0, 0) == 9


def test_morton_spread_all_values(oracle):
    # bit k of a 10-bit value lands on bit 3k: all 1024 values against a bit loop;
    # the three axes use disjoint bit positions, so this covers all 2^30 codes
    L = oracle.lib()
    for v in range(1024):
        naive = sum(((v >> b) & 1) << (3 * b) for b in range(10))
        assert L.or_expand_bits(v) == naive
    assert oracle.morton_encode(1023, 1023, 1023) == (1 << 30) - 1


def test_radix_sort(oracle):
    # tests/main.c:192-233: keys {999,10,500,42} -> {10,42,500,999}, payload follows
    L = oracle.lib()
    codes = np.array([999, 10, 500, 42], np.uint32)
    ids = np.arange(4, dtype=np.uint32)
    v = np.array([9, 1, 5, 4], np.float32)
    x, y, z, m = v.copy(), v.copy(), v.copy(), v.copy()
    L.or_sort(4, codes, ids, x, y, z, m)
    assert codes.tolist() == [10, 42, 500, 999]
    assert m.tolist() == [1.0, 4.0, 5.0, 9.0]
    assert ids.tolist() == [1, 3, 2, 0]


def test_sort_is_stable(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(5)
    codes = rng.integers(0, 16, 5000).astype(np.uint32)
    ids = np.arange(5000, dtype=np.uint32)
    f = np.zeros(5000, np.float32)
    expect = np.argsort(codes, kind="stable").astype(np.uint32)
    L.or_sort(5000, codes, ids, f.copy(), f.copy(), f.copy(), f.copy())
    assert np.array_equal(ids, expect)


def test_octree_insertion(oracle):
    # tests/main.c:125-190: codes {enc(0,0,0), enc(1023,1023,1023)} under a root of
    # half width 10: the child-0 path ends in a leaf (cnt 1, first 0), the
    # child-7 path in a leaf (cnt 1, first 1)
    codes = np.array([oracle.morton_encode(0, 0, 0), oracle.morton_encode(1023, 1023, 1023)],
                     np.uint32)
    t = oracle.Tree(64)
    assert oracle.lib().or_build(2, codes, 0.0, 0.0, 0.0, 10.0, t.st) > 0
    v = t.view()
    for slot, first in ((0, 0), (7, 1)):
        n = 0
        while not v["leaf"][n]:
            assert v["mask"][n] & (1 << slot)
            n = v["child"][n] + bin(v["mask"][n] & ((1 << slot) - 1)).count("1")
        assert v["end"][n] - v["first"][n] == 1 and v["first"][n] == first


def _tick_with_codes(oracle, x, y, z, m, codes, root_hw, theta, centre=0.0):
    n = len(x)
    return oracle.calc_forces(x, y, z, m, np.arange(n, dtype=np.uint32),
                              np.asarray(codes, np.uint32), centre, root_hw, theta)


def test_fmm_upward_pass(oracle):
    # tests/main.c:235-280: (2,3,4) m=2 and (-1,-2,-3) m=3 about the origin:
    # M0=5, MX=1, MY=0, MZ=-1 (tol 1e-4).  Identical codes put both particles in
    # one leaf; the M2M chain shifts the expansion back to the root centre.
    o = _tick_with_codes(oracle, [2, -1], [3, -2], [4, -3], [2, 3], [5, 5], 10.0, THETA_KAT)
    np.testing.assert_allclose(o["tree"]["mp"][0], [5.0, 1.0, 0.0, -1.0], atol=1e-4)


def test_fmm_interaction_pass(oracle):
    # tests/main.c:282-323: A=(-5,-5,-5), B=(5,5,5), hw=1, m0=1, theta=sqrt(12):
    # each local component 0.00192354 (tol 1e-6)
    assert oracle.lib().or_mac_far(-5, -5, -5, 1, 5, 5, 5, 1, THETA_KAT) == 1
    loc = np.zeros(3, np.float32)
    oracle.lib().or_m2l(-5, -5, -5, 5, 5, 5, np.array([1, 0, 0, 0], np.float32), loc)
    np.testing.assert_allclose(loc, [0.00192354] * 3, atol=1e-6)


def test_m2l_dipole_sign(oracle):
    # tests/accuracy.c:74-109: target 0, source (100,0,0), m0=1, mx=1:
    # local_x = 9.799856e-05 +- 5e-9
    assert oracle.lib().or_mac_far(0, 0, 0, 1, 100, 0, 0, 1, 2.0) == 1
    loc = np.zeros(3, np.float32)
    oracle.lib().or_m2l(0, 0, 0, 100, 0, 0, np.array([1, 1, 0, 0], np.float32), loc)
    assert abs(loc[0] - 9.799856e-05) < 5e-9
    # and the closed form of the softened first-order field along d = (100,0,0):
    # E_x = m0 d/r^3 + (p_x/r^3 - 3 (p.d) d/r^5), r^2 = d^2 + 0.1
    r2 = 1e4 + float(np.float32(0.1))
    closed = 100 / r2 ** 1.5 + (1 / r2 ** 1.5 - 3 * 100 * 100 / r2 ** 2.5)
    assert abs(loc[0] - closed) < 1e-10


def test_fmm_downward_and_l2p(oracle):
    # tests/main.c:325-387 and 389-425: one particle m=2.5 in a leaf root; field
    # (5,-3.5,42) gives F=(12.5,-8.75,105); field (2,-1,4) gives (5,-2.5,10)
    L = oracle.lib()
    for field, force in (((5.0, -3.5, 42.0), (12.5, -8.75, 105.0)),
                         ((2.0, -1.0, 4.0), (5.0, -2.5, 10.0))):
        t = oracle.Tree(16)
        L.or_build(1, np.array([oracle.morton_encode(1023, 1023, 1023)], np.uint32),
                   0.0, 0.0, 0.0, 10.0, t.st)
        t.loc[1:4] = field
        f = [np.zeros(1, np.float32) for _ in range(3)]
        L.or_downward(t.st, np.array([2.5], np.float32), *f)
        assert [a[0] for a in f] == list(force)


def test_fmm_p2p_pass(oracle):
    # tests/main.c:427-476: m=2 at 0 and m=3 at x=1 in one leaf:
    # Fx = +-6 * 1.1^-1.5 (tol 1e-4), Fy = Fz = 0
    o = _tick_with_codes(oracle, [0, 1], [0, 0], [0, 0], [2, 3], [0, 0], 10.0, THETA_KAT)
    e = 6.0 * 1.1 ** -1.5
    np.testing.assert_allclose(o["fx"], [e, -e], atol=1e-4)
    assert np.all(o["fy"] == 0) and np.all(o["fz"] == 0)
