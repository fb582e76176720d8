"""The device-resident extensions of the drop-in: ps_calc_forces_async on a
caller's stream and caller-owned device memory (ps_init_ex).

Bars: the asynchronous tick equals the synchronous one bit for bit at the
same lanes per traversal entry (sorted state, forces, tree, pair counts),
and within the parity tolerance at the default lanes; its wait reruns the
traversal after an arena overflow; an engine carved from caller memory
makes no allocation of its own (the library's allocation counter does not
move) and computes what an ordinary engine computes."""
import numpy as np
import pytest

from gpu_helpers import FORCE_RMS_TOL, forces, log, rms_rel
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("codes", "id", "x", "y", "z", "m", "fx", "fy", "fz")


def sync_result(x, y, z, m, trav, lanes, mode=api.MORTON_AVX2, **kw):
    with api.Solver(len(x), theta=2.0, traversal=trav, **kw) as s:
        if lanes is not None:
            api.check("opt", api.ps_set_option(s.ctx, api.OPT_TRAV_LANES, lanes))
        s.load(x, y, z, m)
        s.tick(mode)
        return s.result(), s.stats(), s.tree()


@pytest.mark.parametrize("trav", [api.TRAV_MT, api.TRAV_ST])
def test_async_tick_equals_sync_tick(trav):
    n = 200_003
    x, y, z, m = W.plummer(n, seed=301)
    ref, st_ref, tree_ref = sync_result(x, y, z, m, trav, 8)
    with api.Stream() as stream, api.Solver(n, theta=2.0, traversal=trav) as s:
        api.check("opt", api.ps_set_option(s.ctx, api.OPT_TRAV_LANES, 8))
        s.load(x, y, z, m)
        s.tick_async(stream, api.MORTON_AVX2)
        # a second tick, or a synchronous one, before the wait is refused
        assert api.ps_calc_forces(s.ctx, s.arrs, s.codes, 0.0, 0.0, 0.0, 1.0) == api.PS_EINVAL
        assert api.ps_calc_forces_async(s.ctx, s.arrs, s.codes, 0.0, 0.0, 0.0, 1.0,
                                        stream) == api.PS_EINVAL
        s.wait()
        got = s.result()
        st = s.stats()
        tree = s.tree()
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f
    assert st["n_nodes"] == st_ref["n_nodes"] and st["m2l_pairs"] == st_ref["m2l_pairs"]
    assert st["level_start"] == st_ref["level_start"]
    for k in ("geom", "mp", "loc", "link", "nend", "parent"):
        assert np.array_equal(tree[k], tree_ref[k]), k
    log(f"async = sync: nodes {st['n_nodes']} m2l {st['m2l_pairs']}")


def test_async_tick_default_lanes_and_frames():
    # default lanes (chosen from the node capacity instead of the level
    # sizes): the same sorted state, forces within the parity tolerance;
    # three frames with the device frame step between the ticks
    n = 100_000
    x, y, z, m = W.plummer(n, seed=302)
    ref, _, _ = sync_result(x, y, z, m, api.TRAV_MT, None)
    with api.Stream() as stream, api.Solver(n, theta=2.0) as s:
        s.load(x, y, z, m)
        s.tick_async(stream, api.MORTON_AVX2)
        s.wait()
        got = s.result()
        for f in ("codes", "id", "x", "y", "z", "m"):
            assert np.array_equal(got[f], ref[f]), f
        assert rms_rel(forces(got), forces(ref)) <= FORCE_RMS_TOL
        vel = [api.DeviceArray.from_numpy(np.zeros(n, np.float32)) for _ in range(3)]
        tmp = api.DeviceArray(3 * n, np.float32)
        s.frame_step(vel, tmp, 1e-3)
        for _ in range(2):
            state = [a.numpy() for a in (s.x, s.y, s.z, s.m)]
            s.tick_async(stream, api.MORTON_AVX2)
            s.wait()
            got = s.result()
            want, _, _ = sync_result(*state, api.TRAV_MT, None)
            assert np.array_equal(got["id"], want["id"])
            assert rms_rel(forces(got), forces(want)) <= FORCE_RMS_TOL
            s.frame_step(vel, tmp, 1e-3)


@pytest.mark.parametrize("mode", [api.MORTON_TRUNC, api.MORTON_AVX2])
def test_async_arena_overflow_is_rerun_by_the_wait(mode):
    n = 50_001
    x, y, z, m = W.plummer(n, seed=303)
    ref, st_ref, _ = sync_result(x, y, z, m, api.TRAV_MT, 8, mode)
    with api.Stream() as stream, api.Solver(n, theta=2.0) as s:
        api.check("opt", api.ps_set_option(s.ctx, api.OPT_TRAV_LANES, 8))
        api.check("opt", api.ps_set_option(s.ctx, api.OPT_ARENA_WORDS, 1))
        s.load(x, y, z, m)
        s.tick_async(stream, mode)
        s.wait()
        got = s.result()
        assert s.stats()["m2l_pairs"] == st_ref["m2l_pairs"]
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f


def test_caller_owned_device_memory_allocates_nothing():
    n = 120_000
    x, y, z, m = W.plummer(n, seed=304)
    ref, st_ref, _ = sync_result(x, y, z, m, api.TRAV_MT, 8, stats=True)
    need = api.device_scratch_bytes(n)
    log(f"device scratch for {n} particles: {need} bytes ({need / n:.0f} B per particle)")
    # too small a buffer is refused at ps_init_ex
    small = api.DeviceArray(need // 2, np.uint8)
    rc, ctx = api.ps_init_ex(api.buff_size_for(n), n, 2.0, dev_buff=small,
                             dev_buff_size=small.nbytes)
    assert rc == api.PS_EOOM and ctx is None
    small.free()
    with api.Stream() as stream:
        # caller arrays and the stateless prepare's per-device workspace exist
        # before the count is read
        scratch = api.DeviceArray(need, np.uint8)
        arrays = [api.DeviceArray(n, np.float32) for _ in range(7)]
        idd, codes = api.DeviceArray(n, np.uint32), api.DeviceArray(n, np.uint32)
        arrs = api.make_arrs(*arrays, idd, n)
        for d, a in zip(arrays[:4], (x, y, z, m)):
            d.upload(a)
        idd.upload(np.arange(n, dtype=np.uint32))
        rc, lo, hi, rg = api.ps_prepare_particles(arrs, codes, api.MORTON_AVX2)
        api.check("prepare", rc)
        # the stateless prepare stages host arrays in its own per-device
        # workspace, grown on first use: grow it before counting
        hx, hy, hz, hm = (np.array(a, np.float32) for a in (x, y, z, m))
        hf = [np.zeros(n, np.float32) for _ in range(3)]
        hid, hcodes = np.arange(n, dtype=np.uint32), np.zeros(n, np.uint32)
        harrs = api.make_arrs(hx, hy, hz, hm, *hf, hid, n)
        api.check("prepare", api.ps_prepare_particles(harrs, hcodes, api.MORTON_AVX2)[0])
        before = api.alloc_stats()
        rc, ctx = api.ps_init_ex(api.buff_size_for(n), n, 2.0, dev_buff=scratch,
                                 dev_buff_size=scratch.nbytes)
        api.check("ps_init_ex", rc)
        api.check("lanes", api.ps_set_option(ctx, api.OPT_TRAV_LANES, 8))
        api.check("stats", api.ps_set_option(ctx, api.OPT_STATS, 1))
        c = np.float32(np.float32(lo) + np.float32(rg) / np.float32(2.0))
        hw = np.float32(rg) / np.float32(2.0)
        api.check("calc", api.ps_calc_forces(ctx, arrs, codes, c, c, c, hw))
        got = dict(zip(("x", "y", "z", "m", "fx", "fy", "fz"), (a.numpy() for a in arrays)))
        got.update(id=idd.numpy(), codes=codes.numpy())
        st = api.get_stats(ctx)
        log(f"list words per particle: {st['list_entries'] / n:.1f}")
        # the asynchronous tick of the same state, on the caller's stream
        for d, a in zip(arrays[:4], (x, y, z, m)):
            d.upload(a)
        idd.upload(np.arange(n, dtype=np.uint32))
        api.check("prepare", api.ps_prepare_particles(arrs, codes, api.MORTON_AVX2)[0])
        api.check("async", api.ps_calc_forces_async(ctx, arrs, codes, c, c, c, hw, stream))
        api.check("wait", api.ps_calc_forces_wait(ctx))
        got2 = dict(zip(("x", "y", "z", "m", "fx", "fy", "fz"), (a.numpy() for a in arrays)))
        got2.update(id=idd.numpy(), codes=codes.numpy())
        # host arrays are staged through buffers carved from the same memory
        rc, lo2, _, rg2 = api.ps_prepare_particles(harrs, hcodes, api.MORTON_AVX2)
        api.check("prepare", rc)
        api.check("calc host", api.ps_calc_forces(ctx, harrs, hcodes, c, c, c, hw))
        api.check("destroy", api.ps_destroy(ctx))
        after = api.alloc_stats()
        scratch.free()
    assert after == before, (before, after)
    assert st["p2p_pairs"] == st_ref["p2p_pairs"] and st["m2l_pairs"] == st_ref["m2l_pairs"]
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f
        assert np.array_equal(got2[f], ref[f]), f
    host = dict(x=hx, y=hy, z=hz, m=hm, fx=hf[0], fy=hf[1], fz=hf[2], id=hid, codes=hcodes)
    for f in FIELDS:
        assert np.array_equal(host[f], ref[f]), f


def test_caller_owned_memory_refuses_what_it_does_not_hold():
    # the carve-out holds every buffer of a tick; the correction's training
    # buffers are not among them: with caller memory that call is PS_EOOM,
    # and the library still allocates nothing
    n = 30_000
    x, y, z, m = W.plummer(n, seed=305)
    L = api.hip_lib()
    with api.Solver(n, theta=2.0, device_scratch=True) as s:
        s.load(x, y, z, m)
        s.tick()
        params = api.DeviceArray(10115, np.float32)
        grad = api.DeviceArray(10115, np.float32)
        feat = api.DeviceArray(64 * 24, np.float32)
        target = api.DeviceArray(64 * 3, np.float32)
        before = api.alloc_stats()
        rc = L.ps_hip_mlp_grad(s.ctx.engine, params.ptr, feat.ptr, target.ptr, 64, grad.ptr,
                               None)
        assert rc == api.PS_EOOM
        # the tick's own buffers, including the correction's parameters, fit
        api.check("set_mlp", L.ps_hip_engine_set_mlp(s.ctx.engine, params.ptr))
        api.check("clear_mlp", L.ps_hip_engine_set_mlp(s.ctx.engine, None))
        s.tick()
        assert api.alloc_stats() == before
