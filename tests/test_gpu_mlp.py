"""The learned far-field correction on the device: features, forward,
gradient and Adam kernels against the float64 CPU check (oracle/mlp_ref.py,
itself pinned by finite differences), the fused in-tick application, and a
short training run whose held-out FP64 error must drop below the plain
FMM's (the plain method judges; the gate falls back otherwise)."""
import os
import sys

import numpy as np
import pytest

from gpu_helpers import gpu_tick, log
from polesitter_amd import api, mlp
from polesitter_amd import workloads as W

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_ref as R  # noqa: E402

pytestmark = pytest.mark.gpu


def leaf_of_from_tree(t, n):
    lf = np.zeros(n, np.int64)
    for node in np.nonzero(t["leaf"])[0]:
        lf[t["first"][node]:t["nend"][node]] = node
    return lf


def test_features_match_cpu():
    x, y, z, m = W.plummer(20_000, seed=81)
    L = api.hip_lib()
    n = len(x)
    with api.Solver(n, theta=2.0) as s:
        s.load(x, y, z, m)
        s.tick()
        feat = api.DeviceArray(n * 24, np.float32)
        scale = api.DeviceArray(n, np.float32)
        api.check("features", L.ps_hip_mlp_features(s.ctx.engine, feat.ptr, scale.ptr))
        r = s.result()
        t = s.tree()
        fg = feat.numpy().reshape(n, 24).astype(np.float64)
        sg = scale.numpy().astype(np.float64)
    pos4 = np.stack([r["x"], r["y"], r["z"], r["m"]], 1)
    fr, sr = R.features(pos4, leaf_of_from_tree(t, n), t["geom"], t["mp"], t["loc"],
                        t["parent"])
    assert np.allclose(sg, sr, rtol=2e-6, atol=0), np.abs(sg - sr).max()
    assert np.allclose(fg, fr, rtol=1e-5, atol=1e-5), np.abs(fg - fr).max()


def test_forward_and_gradient_match_cpu():
    rng = np.random.default_rng(82)
    params = R.init_params(seed=83, out_scale=1.0)
    n = 100  # four full blocks of 32 samples and a partial one
    f = rng.normal(size=(n, 24)).astype(np.float32)
    tgt = rng.normal(size=(n, 3)).astype(np.float32)
    t = mlp.Trainer(params)
    try:
        out = t.forward(f)
        ref, _ = R.forward(params.astype(np.float64), f.astype(np.float64))
        assert np.allclose(out, ref, rtol=1e-4, atol=1e-5), np.abs(out - ref).max()
        Fd = api.DeviceArray.from_numpy(f.ravel())
        Td = api.DeviceArray.from_numpy(tgt.ravel())
        loss = t.grad(Fd, Td, n)
        g = t.g.numpy().astype(np.float64)
        lref, gref = R.loss_grad(params.astype(np.float64), f.astype(np.float64),
                                 tgt.astype(np.float64))
        assert abs(loss - lref) <= 1e-4 * lref, (loss, lref)
        rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
        log(f"gradient relative error {rel:.2e}")
        assert rel < 1e-4, rel
        # per layer too: no block of parameters may hide behind the others
        o = 0
        for name, shape in R.SHAPES:
            k = int(np.prod(shape))
            a, b = g[o:o + k], gref[o:o + k]
            assert np.linalg.norm(a - b) <= 1e-4 * max(np.linalg.norm(b), 1e-12), name
            o += k
        # one Adam step against its formula
        p0 = t.params().astype(np.float64)
        t.adam(1e-3)
        m1 = 0.1 * g
        m2 = 0.001 * g * g
        want = p0 - 1e-3 * (m1 / 0.1) / (np.sqrt(m2 / 0.001) + 1e-8)
        assert np.allclose(t.params(), want, rtol=1e-5, atol=1e-7)
    finally:
        t.close()


def test_correction_applied_in_tick():
    x, y, z, m = W.plummer(30_000, seed=84)
    params = R.init_params(seed=85, out_scale=1.0)
    n = len(x)
    L = api.hip_lib()
    plain = gpu_tick(x, y, z, m, tree=False)
    with api.Solver(n, theta=2.0) as s:
        s.load(x, y, z, m)
        api.check("set_mlp", L.ps_hip_engine_set_mlp(s.ctx.engine, params.ctypes.data))
        s.tick()
        feat = api.DeviceArray(n * 24, np.float32)
        scale = api.DeviceArray(n, np.float32)
        api.check("features", L.ps_hip_mlp_features(s.ctx.engine, feat.ptr, scale.ptr))
        corr = s.result()
        fe = feat.numpy().reshape(n, 24)
        sc = scale.numpy().astype(np.float64)
    net, _ = R.forward(params.astype(np.float64), fe.astype(np.float64))
    base = np.stack([plain["fx"], plain["fy"], plain["fz"]], 1).astype(np.float64)
    want = (plain["m"].astype(np.float64) * sc)[:, None] * net
    got = np.stack([corr["fx"], corr["fy"], corr["fz"]], 1).astype(np.float64) - base
    assert np.array_equal(corr["id"], plain["id"])
    # the in-tick net runs on the matrix cores with FP32 operands split into
    # bf16 hi + lo (16-bit mantissas, FP32 accumulation): the correction term
    # itself agrees with the float64 net to 1e-4
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    log(f"correction term relative error {err:.2e}")
    assert err < 1e-4, err
    # particles without a far field (s = 0) are left alone
    assert np.array_equal(got[sc == 0], np.zeros_like(got[sc == 0]))


def test_training_lowers_the_held_out_error():
    # coarse setting: leaf capacity 8; train on an evolving Plummer, validate
    # on another realisation; the correction must beat the plain FMM there
    n = 100_000
    x, y, z, m = W.plummer(n, seed=86)
    v = [np.zeros(n, np.float32)] * 3
    xv, yv, zv, mv = W.plummer(n, seed=87)
    res = mlp.build_correction(x, y, z, m, v, train_ticks=24, val_ticks=6, samples=2048,
                               leaf_cap=8, steps=800, batch=4096, lr=2e-3, seed=1,
                               val_system=(xv, yv, zv, mv, v), log=log)
    log(f"held-out FP64 RMS: plain {res['val_plain']:.5e} corrected {res['val_corrected']:.5e}")
    assert res["enabled"], res
    assert res["val_corrected"] < 0.95 * res["val_plain"], res
