"""The benchmark workloads at full size: exact parity with the oracle at
N = 1,048,576 (Plummer), size-independent properties at the larger sizes,
and the FP64 direct-sum error on 4096 seeded targets matched through id[]
(tests/accuracy.c:505-531)."""
import numpy as np
import pytest

from gpu_helpers import assert_tick_parity, forces, gpu_tick, log, rms_rel
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_direct_f64_device_matches_cpu(oracle):
    x, y, z, m = W.plummer(60_000, seed=61)
    tg = np.array([0, 5, 777, 59_999, 31_000], np.uint32)
    with api.Solver(60_000) as s:
        s.load(x, y, z, m)
        dev = s.direct_f64(tg)
    cpu = oracle.direct_f64(x, y, z, m, tg)
    np.testing.assert_allclose(dev, cpu, rtol=1e-12, atol=0)


def test_plummer_1m_matches_oracle(oracle):
    n = 1 << 20
    x, y, z, m = W.plummer(n, seed=1)
    log("plummer 1M: oracle")
    o = oracle.tick(x, y, z, m, theta=2.0)
    log("plummer 1M: gpu")
    g = gpu_tick(x, y, z, m, theta=2.0)
    f, l = assert_tick_parity(g, o)
    log(f"  nodes {g['tree']['n_nodes']} m2l {g['stats']['m2l_pairs']} "
        f"p2p {g['stats']['p2p_pairs']} force rms {f:.2e}")


def structural_checks(r, tree, n, total_mass):
    codes = r["codes"]
    assert np.all(codes[1:] >= codes[:-1])
    assert np.array_equal(np.sort(r["id"]), np.arange(n, dtype=np.uint32))
    lf = tree["leaf"]
    first = tree["first"][lf].astype(np.int64)
    end = tree["nend"][lf].astype(np.int64)
    order = np.argsort(first)
    assert first[order][0] == 0 and end[order][-1] == n
    assert np.array_equal(first[order][1:], end[order][:-1])  # leaves tile [0, n)
    assert abs(tree["mp"][0, 0] - total_mass) < 1e-4 * total_mass


@pytest.mark.parametrize("n", [1 << 22])
def test_plummer_4m_properties_and_fp64_error(n):
    x, y, z, m = W.plummer(n, seed=2)
    log(f"plummer {n}")
    with api.Solver(n, theta=2.0, stats=True) as s:
        s.load(x, y, z, m)
        s.tick()
        r = s.result()
        structural_checks(r, s.tree(), n, float(m.astype(np.float64).sum()))
        tg = W.sample_targets(n, 4096, seed=4096)
        inv = np.empty(n, np.uint32)
        inv[r["id"]] = np.arange(n, dtype=np.uint32)
        pos = inv[tg]
        ref = s.direct_f64(pos)
    f = np.stack([r["fx"], r["fy"], r["fz"]], 1)[pos].astype(np.float64)
    rms = rms_rel(f, ref)
    log(f"  FP64 RMS error on 4096 targets: {rms:.3e}")
    # the first-order FMM's own truncation error at theta=2 for a deep Plummer
    # tree (7.4e-2 measured here); the GPU matches the oracle exactly at 1M,
    # so this bound only guards against a broken far field
    assert rms < 0.15, rms


def test_ticks_with_euler_stay_consistent(oracle):
    # 10 ticks with semi-implicit Euler dt=1e-3: every tick re-sorts and rebuilds;
    # tick k's forces match the oracle on the same state
    n = 1 << 18
    x, y, z, m = W.plummer(n, seed=3)
    with api.Solver(n, theta=2.0) as s:
        s.load(x, y, z, m)
        vel = [api.DeviceArray.from_numpy(np.zeros(n, np.float32)) for _ in range(3)]
        tmp = api.DeviceArray(3 * n, np.float32)
        L = api.hip_lib()
        for t in range(10):
            state = (s.x.numpy(), s.y.numpy(), s.z.numpy(), s.m.numpy())
            s.reset_ids()
            s.tick()
            if t in (0, 9):
                o = oracle.tick(*state, theta=2.0)
                assert rms_rel(s.forces().astype(np.float64), forces(o)) <= 1e-5
                assert np.array_equal(s.id.numpy(), o["id"])
            api.check("gather", L.ps_hip_step_gather(s.ctx.engine, s.id.ptr, vel[0].ptr,
                                                    vel[1].ptr, vel[2].ptr, tmp.ptr, n))
            api.check("euler", L.ps_hip_step_euler(s.ctx.engine, s.x.ptr, s.y.ptr, s.z.ptr,
                                                  vel[0].ptr, vel[1].ptr, vel[2].ptr, s.fx.ptr,
                                                  s.fy.ptr, s.fz.ptr, s.m.ptr, 1e-3, n))


@pytest.mark.parametrize("n", [(1 << 22) + 12_345])
def test_large_sort_is_the_stable_order(n):
    # at 4M keys (many tiles, a partial last tile) the permutation is the
    # stable one (equal codes keep their input order) and the payload follows
    # it
    x, y, z, m = W.plummer(n, seed=5)
    log(f"sort {n}")
    g = gpu_tick(x, y, z, m, theta=2.0, tree=False)
    perm = np.argsort(g["prep_codes"], kind="stable")
    assert np.array_equal(g["id"], perm.astype(np.uint32))
    assert np.array_equal(g["codes"], g["prep_codes"][perm])
    for k, a in (("x", x), ("y", y), ("z", z), ("m", m)):
        assert np.array_equal(g[k], np.asarray(a, np.float32)[perm]), k
