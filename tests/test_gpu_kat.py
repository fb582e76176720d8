"""The reference's known-answer tests (tests/main.c, tests/accuracy.c) run on
the gfx950 kernels: single-operator entry points use the tick's own device
functions; the staged entry points run the tick's kernels one pass at a
time, as the reference's white-box tests call its static passes."""
import ctypes as C

import numpy as np
import pytest

from gpu_helpers import log
from polesitter_amd import api

pytestmark = pytest.mark.gpu

THETA_KAT = 3.4641  # sqrt(12), tests/main.c:11


def enc(x, y, z):
    s = lambda v: sum(((v >> b) & 1) << (3 * b) for b in range(10))
    return s(x) | (s(y) << 1) | (s(z) << 2)


def test_device_present():
    name = api.device_name()
    log(f"device: {name}")
    assert "gfx950" in name


def test_morton_exhaustive_2_30():
    # every (x,y,z) in [0,1024)^3 against a bit loop; the interleave is a
    # bijection of [0, 2^30), so the codes sum to 2^30 (2^30 - 1) / 2
    out = (C.c_uint64 * 2)()
    api.check("op_morton", api.hip_lib().ps_hip_op_morton_exhaustive(out))
    assert out[0] == 0
    assert out[1] == (1 << 30) * ((1 << 30) - 1) // 2


def op_m2l(t, s, mp, theta):
    # keep every array alive across the call: a temporary's .ctypes.data
    # dangles as soon as the temporary is released
    tg, sg, smp = (np.ascontiguousarray(v, np.float32) for v in (t, s, mp))
    out = np.zeros(4, np.float32)
    api.check("op_m2l", api.hip_lib().ps_hip_op_m2l(tg.ctypes.data, sg.ctypes.data,
                                                   smp.ctypes.data, theta, out.ctypes.data))
    return out


def test_fmm_interaction_pass():
    # tests/main.c:282-323
    out = op_m2l([-5, -5, -5, 1], [5, 5, 5, 1], [1, 0, 0, 0], THETA_KAT)
    assert out[0] == 1.0
    np.testing.assert_allclose(out[1:], [0.00192354] * 3, atol=1e-6)


def test_m2l_dipole_sign():
    # tests/accuracy.c:74-109
    out = op_m2l([0, 0, 0, 1], [100, 0, 0, 1], [1, 1, 0, 0], 2.0)
    assert out[0] == 1.0 and abs(out[1] - 9.799856e-05) < 5e-9


def test_mac_tie_is_near():
    # theta=2, equal cells at offset (4h,0,0): dist^2 == theta^2 (2h)^2 exactly,
    # which is NOT far (strict >, H:1064)
    out = op_m2l([0, 0, 0, 0.5], [2, 0, 0, 0.5], [1, 0, 0, 0], 2.0)
    assert out[0] == 0.0
    out = op_m2l([0, 0, 0, 0.5], [2.0000002, 0, 0, 0.5], [1, 0, 0, 0], 2.0)
    assert out[0] == 1.0


class Staged:
    """Device arrays + one engine, driven stage by stage."""

    def __init__(self, x, y, z, m, codes, root, theta=THETA_KAT):
        n = len(x)
        self.n = n
        self.s = api.Solver(n, theta=theta)
        self.s.load(x, y, z, m)
        self.s.codes.upload(np.asarray(codes, np.uint32))
        self.e = self.s.ctx.engine
        api.check("bind", api.hip_lib().ps_hip_debug_bind(
            self.e, self.s.x.ptr, self.s.y.ptr, self.s.z.ptr, self.s.m.ptr, self.s.fx.ptr,
            self.s.fy.ptr, self.s.fz.ptr, self.s.id.ptr, n, self.s.codes.ptr, *root))

    def stage(self, *ks):
        for k in ks:
            api.check(f"stage {k}", api.hip_lib().ps_hip_debug_stage(self.e, k))

    def tree(self, cap=4096):
        return api.export_tree(self.s.ctx, cap)

    def close(self):
        self.s.close()


def test_radix_sort():
    # tests/main.c:192-233
    v = [9.0, 1.0, 5.0, 4.0]
    st = Staged(v, v, v, v, [999, 10, 500, 42], (0.0, 0.0, 0.0, 10.0))
    st.stage(1)
    r = st.s.result()
    assert r["codes"].tolist() == [10, 42, 500, 999]
    assert r["m"].tolist() == [1.0, 4.0, 5.0, 9.0]
    assert r["id"].tolist() == [1, 3, 2, 0]
    st.close()


def test_radix_sort_stable_large():
    rng = np.random.default_rng(7)
    n = 300_000
    codes = rng.integers(0, 1 << 12, n).astype(np.uint32) << np.uint32(18)  # many ties
    f = rng.random(n).astype(np.float32)
    st = Staged(f, f, f, f, codes, (0.0, 0.0, 0.0, 1.0))
    st.stage(1)
    r = st.s.result()
    perm = np.argsort(codes, kind="stable")
    assert np.array_equal(r["id"], perm.astype(np.uint32))
    assert np.array_equal(r["codes"], codes[perm]) and np.array_equal(r["x"], f[perm])
    st.close()


def walk(t, slot):
    n = 0
    while not t["leaf"][n]:
        mask = int(t["mask"][n])
        assert mask & (1 << slot), "missing child on the path"
        n = int(t["link"][n, 0]) + bin(mask & ((1 << slot) - 1)).count("1")
    return n


def test_octree_insertion():
    # tests/main.c:125-190
    st = Staged([0, 0], [0, 0], [0, 0], [1, 1], [enc(0, 0, 0), enc(1023, 1023, 1023)],
                (0.0, 0.0, 0.0, 10.0))
    st.stage(2)
    t = st.tree()
    for slot, first in ((0, 0), (7, 1)):
        n = walk(t, slot)
        assert t["nend"][n] - t["first"][n] == 1 and t["first"][n] == first
    st.close()


def test_fmm_upward_pass():
    # tests/main.c:235-280: M0=5, MX=1, MY=0, MZ=-1 about the origin
    st = Staged([2, -1], [3, -2], [4, -3], [2, 3], [5, 5], (0.0, 0.0, 0.0, 10.0))
    st.stage(2, 3)
    np.testing.assert_allclose(st.tree()["mp"][0], [5, 1, 0, -1], atol=1e-4)
    st.close()


@pytest.mark.parametrize("field,force", [((5.0, -3.5, 42.0), (12.5, -8.75, 105.0)),
                                         ((2.0, -1.0, 4.0), (5.0, -2.5, 10.0))])
def test_fmm_downward_and_l2p(field, force):
    # tests/main.c:325-387 and 389-425
    st = Staged([0], [0], [0], [2.5], [enc(1023, 1023, 1023)], (0.0, 0.0, 0.0, 10.0))
    st.stage(2)
    f3 = np.asarray(field, np.float32)
    api.check("set_local", api.hip_lib().ps_hip_debug_set_local(st.e, 0, f3.ctypes.data))
    st.stage(5)
    r = st.s.result()
    assert (r["fx"][0], r["fy"][0], r["fz"][0]) == force
    assert tuple(st.tree()["loc"][0, 1:]) == field
    st.close()


def test_fmm_downward_l2l_chain():
    # L2L down a chain: a leaf several levels below the root sees the root field
    st = Staged([0, 0], [0, 0], [0, 0], [2.5, 1.0], [enc(1023, 1023, 1023)] * 2,
                (0.0, 0.0, 0.0, 10.0))
    st.stage(2)
    f3 = np.asarray([5.0, -3.5, 42.0], np.float32)
    api.check("set_local", api.hip_lib().ps_hip_debug_set_local(st.e, 0, f3.ctypes.data))
    st.stage(5)
    t = st.tree()
    leaf = walk(t, 7)
    assert tuple(t["loc"][leaf, 1:]) == (5.0, -3.5, 42.0)
    r = st.s.result()
    assert (r["fx"][0], r["fy"][0], r["fz"][0]) == (12.5, -8.75, 105.0)
    st.close()


def test_fmm_p2p_pass():
    # tests/main.c:427-476: Fx = +-6 * 1.1^-1.5 (tol 1e-4)
    st = Staged([0, 1], [0, 0], [0, 0], [2, 3], [0, 0], (0.0, 0.0, 0.0, 10.0))
    st.stage(2, 3, 4, 5, 6)
    r = st.s.result()
    e = 6.0 * 1.1 ** -1.5
    np.testing.assert_allclose(r["fx"], [e, -e], atol=1e-4)
    assert np.all(r["fy"] == 0) and np.all(r["fz"] == 0)
    st.close()


def morton_codes(x, y, z, hw):
    # 10-bit quantisation of [-hw, hw)^3 about the origin, interleaved x, y, z
    out = np.zeros(len(x), np.uint32)
    for k, v in enumerate((x, y, z)):
        q = np.clip(np.floor((v.astype(np.float64) + hw) / (2 * hw) * 1024), 0, 1023)
        q = q.astype(np.uint32)
        for b in range(10):
            out |= ((q >> np.uint32(b)) & np.uint32(1)) << np.uint32(3 * b + k)
    return out


@pytest.mark.parametrize("seed", [3, 11])
def test_tick_tail_matches_the_staged_passes(seed):
    # the tick's L2L (the leaves' summed locals in one launch, the node locals
    # summed when read) against the level-by-level downward pass: the same
    # forces and the same node locals, bit for bit
    from polesitter_amd import workloads as W
    x, y, z, m = W.two_plummers(60_000, seed=seed)[:4]
    hw = float(max(np.abs(x).max(), np.abs(y).max(), np.abs(z).max())) * 1.01
    codes = morton_codes(x, y, z, hw)
    root = (0.0, 0.0, 0.0, hw)
    out = []
    for stages in ((1, 2, 3, 4, 5, 6), (1, 2, 3, 7)):
        log(f"  seed {seed} stages {stages}")
        st = Staged(x, y, z, m, codes, root, theta=2.0)
        st.stage(*stages)
        r = st.s.result()
        t1 = st.tree(4 * len(x))
        t2 = st.tree(4 * len(x))  # a second read sums nothing again
        assert np.array_equal(t1["loc"].view(np.uint32), t2["loc"].view(np.uint32))
        out.append((r, t1))
        st.close()
    (ra, ta), (rb, tb) = out
    for k in ("fx", "fy", "fz"):
        assert np.array_equal(ra[k].view(np.uint32), rb[k].view(np.uint32)), k
    assert np.array_equal(ta["loc"].view(np.uint32), tb["loc"].view(np.uint32))
    assert np.abs(ta["loc"][:, 1:]).max() > 0  # the far field is there
