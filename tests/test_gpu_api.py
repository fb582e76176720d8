"""Error behaviour of the drop-in API on the device (result codes of
H:294-300), the caller's frame step, and the README loop as a C program
linked against the library."""
import os
import subprocess

import numpy as np
import pytest

from gpu_helpers import forces, gpu_tick, log, rms_rel
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_init_buffer_contract():
    # radix scratch of 24 B per particle must fit after the context (H:1717-1723)
    rc, ctx = api.ps_init(4096, 1000)
    assert rc == api.PS_EOOM and ctx is None
    rc, ctx = api.ps_init(1 << 20, 1000)
    assert rc == api.PS_OK
    assert api.ps_destroy(ctx) == api.PS_OK


def test_calc_forces_errors():
    n = 1000
    x, y, z, m = W.uniform_cube(n, seed=51)
    with api.Solver(n) as s:
        s.load(x, y, z, m)
        rc, lo, hi, rg = api.ps_prepare_particles(s.arrs, s.codes)
        assert rc == api.PS_OK
        # cnt beyond max_particles: the reference overruns its scratch; here EINVAL
        big = api.make_arrs(s.x, s.y, s.z, s.m, s.fx, s.fy, s.fz, s.id, n + 1)
        assert api.ps_calc_forces(s.ctx, big, s.codes, 0, 0, 0, 1) == api.PS_EINVAL
        # no codes
        assert api.ps_calc_forces(s.ctx, s.arrs, None, 0, 0, 0, 1) == api.PS_EINVAL
        # cnt == 0 is a no-op (H:1592)
        empty = api.make_arrs(s.x, s.y, s.z, s.m, s.fx, s.fy, s.fz, s.id, 0)
        assert api.ps_calc_forces(s.ctx, empty, s.codes, 0, 0, 0, 1) == api.PS_OK
        # codes wider than 30 bits are not Morton codes of ps_prepare_particles
        c = s.codes.numpy()
        c[7] |= np.uint32(1 << 31)
        s.codes.upload(c)
        assert api.ps_calc_forces(s.ctx, s.arrs, s.codes, 0, 0, 0, 1) == api.PS_EINVAL


def test_prepare_errors():
    arrs = api.make_arrs(None, None, None, None, None, None, None, None, 10)
    rc, *_ = api.ps_prepare_particles(arrs, None)
    assert rc == api.PS_EINVAL
    x = api.DeviceArray(4, np.float32)
    c = api.DeviceArray(4, np.uint32)
    arrs = api.make_arrs(x, x, x, None, None, None, None, None, 0)
    rc, *_ = api.ps_prepare_particles(arrs, c)
    assert rc == api.PS_EINVAL  # cnt == 0, H:1914


def test_node_capacity_exceeded_is_reported():
    # a buffer that passes ps_init but holds too few nodes: the reference
    # silently truncates the tree (H:930-933); the drop-in returns PS_EOOM
    n = 5000
    x, y, z, m = W.uniform_cube(n, seed=52)
    size = 64 + 24 * n + 80 * 100  # room for ~100 nodes
    with api.Solver(n, buff_size=size) as s:
        s.load(x, y, z, m)
        rc, lo, hi, rg = api.ps_prepare_particles(s.arrs, s.codes)
        c = np.float32(lo + rg / 2)
        assert api.ps_calc_forces(s.ctx, s.arrs, s.codes, c, c, c, rg / 2) == api.PS_EOOM


def test_reuse_context_across_sizes(oracle):
    # one context, ticks of different N in a row (buffers are reused / grown)
    with api.Solver(60_000, theta=2.0, stats=True) as s:
        for n, seed in ((60_000, 53), (5_000, 54), (60_000, 55), (1, 56)):
            x, y, z, m = W.plummer(n, seed=seed)
            arrs = api.make_arrs(s.x, s.y, s.z, s.m, s.fx, s.fy, s.fz, s.id, n)
            for d, a in ((s.x, x), (s.y, y), (s.z, z), (s.m, m)):
                full = np.zeros(60_000, np.float32)
                full[:n] = a
                d.upload(full)
            s.reset_ids()
            rc, lo, hi, rg = api.ps_prepare_particles(arrs, s.codes)
            assert rc == api.PS_OK
            c = np.float32(np.float32(lo) + np.float32(rg) * np.float32(0.5))
            hw = np.float32(rg) * np.float32(0.5)
            assert api.ps_calc_forces(s.ctx, arrs, s.codes, c, c, c, hw) == api.PS_OK
            o = oracle.tick(x, y, z, m, theta=2.0)
            f = np.stack([a.numpy()[:n] for a in (s.fx, s.fy, s.fz)], 1).astype(np.float64)
            assert rms_rel(f, forces(o)) <= 1e-5
            assert np.array_equal(s.id.numpy()[:n], o["id"])


def test_frame_step_on_device(oracle):
    # README.md:94-161 on the device: reset ids, tick, gather velocities by id,
    # semi-implicit Euler -- against the same loop in numpy + the oracle
    n = 20_000
    x, y, z, m, vx, vy, vz = W.two_plummers(n, seed=57, sep=2.0)
    dt = np.float32(1e-3)
    hx, hy, hz, hm = x.copy(), y.copy(), z.copy(), m.copy()
    hv = [vx.copy(), vy.copy(), vz.copy()]
    with api.Solver(n, theta=2.0) as s:
        s.load(x, y, z, m)
        dv = [api.DeviceArray.from_numpy(v) for v in (vx, vy, vz)]
        tmp = api.DeviceArray(3 * n, np.float32)
        L = api.hip_lib()
        for tick in range(3):
            s.reset_ids()
            s.tick()
            api.check("gather", L.ps_hip_step_gather(s.ctx.engine, s.id.ptr, dv[0].ptr,
                                                    dv[1].ptr, dv[2].ptr, tmp.ptr, n))
            api.check("euler", L.ps_hip_step_euler(s.ctx.engine, s.x.ptr, s.y.ptr, s.z.ptr,
                                                  dv[0].ptr, dv[1].ptr, dv[2].ptr, s.fx.ptr,
                                                  s.fy.ptr, s.fz.ptr, s.m.ptr, float(dt), n))
            # the same frame on the host with the oracle's forces
            o = oracle.tick(hx, hy, hz, hm, theta=2.0)
            hx, hy, hz, hm = o["x"], o["y"], o["z"], o["m"]
            hv = [v[o["id"]] for v in hv]
            assert np.array_equal(s.id.numpy(), o["id"])
            fo = forces(o)
            assert rms_rel(s.forces().astype(np.float64), fo) <= 1e-5
            # continue both runs from the device state so roundoff cannot drift
            hx, hy, hz = s.x.numpy(), s.y.numpy(), s.z.numpy()
            for k in range(3):
                v_dev = dv[k].numpy()
                f_o = fo[:, k].astype(np.float32)
                v_host = hv[k] + (f_o / hm) * dt
                np.testing.assert_allclose(v_dev, v_host, rtol=1e-4, atol=1e-6)
                hv[k] = v_dev


def test_fused_frame_step_equals_separate_steps():
    # ps_hip_step_frame = gather + Euler + reset ids, bit for bit, massless
    # particles included (they keep x and the gathered velocity)
    n = 30_000
    x, y, z, m, vx, vy, vz = W.two_plummers(n, seed=58, sep=2.0)
    m = m.copy()
    m[::97] = 0.0
    dt = 1e-3
    L = api.hip_lib()
    outs = []
    for fused in (False, True):
        with api.Solver(n, theta=2.0) as s:
            s.load(x, y, z, m)
            dv = [api.DeviceArray.from_numpy(v) for v in (vx, vy, vz)]
            tmp = api.DeviceArray(3 * n, np.float32)
            for _ in range(3):
                s.tick()
                if fused:
                    s.frame_step(dv, tmp, dt)
                else:
                    api.check("gather", L.ps_hip_step_gather(s.ctx.engine, s.id.ptr, dv[0].ptr,
                                                            dv[1].ptr, dv[2].ptr, tmp.ptr, n))
                    api.check("euler", L.ps_hip_step_euler(
                        s.ctx.engine, s.x.ptr, s.y.ptr, s.z.ptr, dv[0].ptr, dv[1].ptr,
                        dv[2].ptr, s.fx.ptr, s.fy.ptr, s.fz.ptr, s.m.ptr, dt, n))
                    s.reset_ids()
            r = s.result()
            outs.append([r[k] for k in ("x", "y", "z", "m", "fx", "id")] +
                        [v.numpy() for v in dv])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert np.array_equal(outs[1][5], np.arange(n, dtype=np.uint32))


def test_dropin_c_program(oracle, tmp_path):
    # the reference README's frame loop compiled against include/polesitter.h
    exe = os.path.join(api.LIBDIR, "dropin_tick")
    out = tmp_path / "tick.bin"
    n = 4096
    r = subprocess.run([exe, str(n), "2.0", str(out), "3"], capture_output=True, text=True,
                       timeout=120)
    log(r.stdout.strip())
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint32).reshape(9, n)
    f32 = raw[:7].view(np.float32)
    x, y, z, m = W.ref_fill_uniform(n)
    o = oracle.tick(x, y, z, m, theta=2.0)
    assert np.array_equal(raw[7], o["id"]) and np.array_equal(raw[8], o["codes"])
    assert np.array_equal(f32[0], o["x"]) and np.array_equal(f32[3], o["m"])
    f = np.stack(f32[4:7], 1).astype(np.float64)
    assert rms_rel(f, forces(o)) <= 1e-5


def test_stats_modes_time_what_they_say():
    # 2: every stage event; 3: the traversal and the P2P stage only; 0: none
    x, y, z, m = W.plummer(50_000, seed=41)
    with api.Solver(len(x), theta=2.0) as s:
        got = {}
        for mode in (2, 3, 0):
            api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_STATS, mode))
            s.load(x, y, z, m)
            s.tick(api.MORTON_TRUNC)
            got[mode] = s.stats()
        assert api.ps_set_option(s.ctx, api.OPT_STATS, 4) != api.PS_OK
    full, light, off = got[2], got[3], got[0]
    for k in ("ms_sort", "ms_build", "ms_upward", "ms_traverse", "ms_p2p", "ms_total",
              "ms_trav_steps"):
        assert full[k] > 0.0, k
    assert light["ms_trav_steps"] > 0.0 and light["ms_p2p"] > 0.0
    for k in ("ms_sort", "ms_build", "ms_upward", "ms_total"):
        assert light[k] == 0.0, k
    for k in ("ms_sort", "ms_p2p", "ms_total", "ms_trav_steps"):
        assert off[k] == 0.0, k
    assert full["m2l_pairs"] == light["m2l_pairs"] == off["m2l_pairs"] > 0


def test_mlp_entry_points_reject_host_pointers():
    # the training kernels read and write their arguments directly; a host
    # array would fault the device, so each entry point checks every pointer
    # first and returns PS_EINVAL (no kernel is launched)
    L = api.hip_lib()
    P = 10115
    with api.Solver(64) as s:
        e = s.ctx.engine
        dp, dg, d1, d2 = (api.DeviceArray(P, np.float32) for _ in range(4))
        feat = api.DeviceArray(4 * 24, np.float32)
        out = api.DeviceArray(4 * 3, np.float32)
        hp = np.zeros(P, np.float32)
        hf = np.zeros(4 * 24, np.float32)
        ho = np.zeros(4 * 3, np.float32)
        assert L.ps_hip_mlp_forward(e, hp.ctypes.data, feat.ptr, 4, out.ptr) == api.PS_EINVAL
        assert L.ps_hip_mlp_forward(e, dp.ptr, hf.ctypes.data, 4, out.ptr) == api.PS_EINVAL
        assert L.ps_hip_mlp_forward(e, dp.ptr, feat.ptr, 4, ho.ctypes.data) == api.PS_EINVAL
        assert L.ps_hip_mlp_grad(e, dp.ptr, feat.ptr, ho.ctypes.data, 4, dg.ptr,
                                 None) == api.PS_EINVAL
        assert L.ps_hip_mlp_grad(e, dp.ptr, feat.ptr, out.ptr, 4, hp.ctypes.data,
                                 None) == api.PS_EINVAL
        assert L.ps_hip_mlp_adam(e, dp.ptr, dg.ptr, hp.ctypes.data, d2.ptr, 1e-3, 0.9, 0.999,
                                 1e-8, 1) == api.PS_EINVAL
        assert L.ps_hip_mlp_adam(e, hp.ctypes.data, dg.ptr, d1.ptr, d2.ptr, 1e-3, 0.9, 0.999,
                                 1e-8, 1) == api.PS_EINVAL
        # the device arrays themselves are accepted
        api.check("forward", L.ps_hip_mlp_forward(e, dp.ptr, feat.ptr, 4, out.ptr))
