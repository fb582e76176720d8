"""Training and validation of the learned far-field correction.

The correction (kernels in csrc/ps_mlp.hip) predicts, per particle, the
far-field error of the first-order FMM in units of the leaf's field
magnitude s = |E_L|:  a_exact ~ a_fmm + s * net(F).  The pipeline:

1. ``collect``: ticks of a workload at the (coarse) setting to be corrected,
   with the correction off; for seeded targets of every tick, the features
   F, the scale s, the FMM acceleration and the FP64 direct-sum acceleration
   (tests/accuracy.c:42-72, on the device).  Between ticks the caller's frame
   step (id-gather, semi-implicit Euler) moves the particles.
2. ``train``: minibatch Adam on the mean squared error of the normalised
   residual t = (a_exact - a_fmm) / s, with the project's own forward and
   backward kernels (``ps_hip_mlp_grad``, ``ps_hip_mlp_adam``).
3. ``validate``: on ticks held out from training, the FP64 RMS force error
   with and without the correction.  The plain FMM is the judge: the
   correction is enabled only if it lowers that error (``enabled``).

All GPU work goes through the C-ABI (``api``); nothing here runs on the
CPU except the bookkeeping.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import api
from . import workloads as W

N_IN = 24
N_PARAMS = 10115


def init_params(seed: int = 0, out_scale: float = 0.1) -> np.ndarray:
    """He-style initialisation with a small output layer (the untrained
    correction starts near zero).  Packed layout as in the kernels."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in, last in ((64, 24, False), (64, 64, False), (64, 64, False),
                                  (3, 64, True)):
        w = rng.normal(0.0, np.sqrt(2.0 / fan_in), (fan_out, fan_in))
        if last:
            w *= out_scale
        parts += [w.ravel(), np.zeros(fan_out)]
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == N_PARAMS
    return p


def _log(log, msg):
    if log:
        log(msg)


def collect(x, y, z, m, v, ticks: int, samples: int = 4096, dt: float = 1e-3,
            theta: float = 2.0, leaf_cap: int = 1, mode: int = api.MORTON_AVX2,
            seed: int = 0, log=None) -> dict:
    """Residual samples of `ticks` ticks of one evolving system (the caller's
    frame step with time step dt between ticks)."""
    n = len(x)
    L = api.hip_lib()
    F, S, AF, AX, M = [], [], [], [], []
    with api.Solver(n, theta=theta) as s:
        api.check("ps_set_option", api.ps_set_option(s.ctx, api.OPT_LEAF_CAP, leaf_cap))
        s.load(x, y, z, m)
        vel = [api.DeviceArray.from_numpy(np.asarray(c, np.float32)) for c in v]
        tmp = api.DeviceArray(3 * n, np.float32)
        rows = min(samples, n)
        idx = api.DeviceArray(rows, np.uint32)
        feat = api.DeviceArray(rows * N_IN, np.float32)
        scale = api.DeviceArray(rows, np.float32)
        for k in range(ticks):  # ids are 0..n-1 here (load, then every frame step)
            s.tick(mode)
            ids = s.id.numpy()
            tg = W.sample_targets(n, rows, seed=seed * 100003 + k)
            inv = np.empty(n, np.int64)
            inv[ids] = np.arange(n)
            pos = np.ascontiguousarray(inv[tg], np.uint32)
            idx.upload(pos)
            api.check("ps_hip_mlp_features_at", L.ps_hip_mlp_features_at(
                s.ctx.engine, idx.ptr, rows, feat.ptr, scale.ptr))
            fx = s.direct_f64(pos)
            mass = s.m.numpy()[pos].astype(np.float64)
            ok = mass > 0
            fa = np.stack([s.fx.numpy()[pos], s.fy.numpy()[pos], s.fz.numpy()[pos]],
                          1).astype(np.float64)
            fe = feat.numpy().reshape(rows, N_IN)
            sc = scale.numpy().astype(np.float64)
            F.append(fe[ok])
            S.append(sc[ok])
            AF.append(fa[ok] / mass[ok, None])
            AX.append(fx[ok] / mass[ok, None])
            M.append(mass[ok])
            # the caller's frame step (README.md:94-161) on the device
            s.frame_step(vel, tmp, dt)
            if (k + 1) % 20 == 0 or k + 1 == ticks:
                _log(log, f"mlp: collected tick {k + 1}/{ticks}")
    S_ = np.concatenate(S)
    AF_ = np.concatenate(AF)
    AX_ = np.concatenate(AX)
    keep = S_ > 0
    T = np.where(keep[:, None], (AX_ - AF_) / np.where(keep, S_, 1.0)[:, None], 0.0)
    return dict(F=np.concatenate(F).astype(np.float32), T=T.astype(np.float32), s=S_,
                a_fmm=AF_, a_exact=AX_, m=np.concatenate(M))


def force_rms(d: dict, pred=None) -> float:
    """FP64 RMS relative force error (tests/accuracy.c:505-531) of the FMM
    accelerations, corrected by s * pred when pred is given"""
    a = d["a_fmm"] if pred is None else d["a_fmm"] + d["s"][:, None] * pred
    m = d["m"][:, None]
    e = ((a - d["a_exact"]) * m) ** 2
    r = (d["a_exact"] * m) ** 2
    return float(np.sqrt(e.sum() / r.sum()))


class Trainer:
    """Device state of one training run (parameters, Adam moments, data)."""

    def __init__(self, params: np.ndarray):
        self.L = api.hip_lib()
        self.solver = api.Solver(64)  # an engine: its stream and scratch
        self.eng = self.solver.ctx.engine
        self.p = api.DeviceArray.from_numpy(np.asarray(params, np.float32))
        self.g = api.DeviceArray(N_PARAMS, np.float32)
        self.m1 = api.DeviceArray.from_numpy(np.zeros(N_PARAMS, np.float32))
        self.m2 = api.DeviceArray.from_numpy(np.zeros(N_PARAMS, np.float32))
        self.step = 0

    def params(self) -> np.ndarray:
        return self.p.numpy()

    def grad(self, Fd, Td, n, off=0):
        loss = C.c_float()
        api.check("ps_hip_mlp_grad", self.L.ps_hip_mlp_grad(
            self.eng, self.p.ptr, C.c_void_p(Fd.ptr + 4 * N_IN * off),
            C.c_void_p(Td.ptr + 12 * off), n, self.g.ptr, C.byref(loss)))
        return float(loss.value)

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        api.check("ps_hip_mlp_adam", self.L.ps_hip_mlp_adam(
            self.eng, self.p.ptr, self.g.ptr, self.m1.ptr, self.m2.ptr, lr, b1, b2, eps,
            self.step))

    def forward(self, F: np.ndarray) -> np.ndarray:
        n = F.shape[0]
        Fd = api.DeviceArray.from_numpy(np.ascontiguousarray(F, np.float32).ravel())
        out = api.DeviceArray(3 * n, np.float32)
        api.check("ps_hip_mlp_forward", self.L.ps_hip_mlp_forward(self.eng, self.p.ptr, Fd.ptr,
                                                                  n, out.ptr))
        return out.numpy().reshape(n, 3).astype(np.float64)

    def close(self):
        self.solver.close()


def train(train_set: dict, steps: int = 3000, batch: int = 8192, lr: float = 1e-3,
          seed: int = 0, log=None, params=None):
    """Adam on the normalised residuals; returns (params, trainer, history)."""
    rng = np.random.default_rng(seed)
    t = Trainer(init_params(seed) if params is None else params)
    F, T = train_set["F"], train_set["T"]
    n = F.shape[0]
    batch = min(batch, n)
    hist = []
    Fd = Td = None
    per_epoch = max(1, n // batch)
    for k in range(steps):
        if k % per_epoch == 0:  # a new shuffle, uploaded once per epoch
            perm = rng.permutation(n)
            Fd = api.DeviceArray.from_numpy(np.ascontiguousarray(F[perm]).ravel())
            Td = api.DeviceArray.from_numpy(np.ascontiguousarray(T[perm]).ravel())
        off = (k % per_epoch) * batch
        loss = t.grad(Fd, Td, batch, off)
        t.adam(lr * (0.1 ** (k / max(steps, 1))))  # decays to lr/10 over the run
        hist.append(loss)
        if log and ((k + 1) % 500 == 0 or k + 1 == steps):
            _log(log, f"mlp: step {k + 1}/{steps} loss {np.mean(hist[-100:]):.4e}")
    return t.params(), t, hist


def build_correction(x, y, z, m, v, train_ticks=200, val_ticks=20, samples=4096, dt=1e-3,
                     theta=2.0, leaf_cap=1, steps=3000, batch=8192, lr=1e-3, seed=0,
                     val_system=None, log=None) -> dict:
    """Collect, train and validate.  val_system = (x, y, z, m, v) of another
    realisation for the held-out ticks (default: the same system continued
    after the training ticks).  Returns a dict with the parameters, the
    validation errors and `enabled` (the fallback decision)."""
    tr = collect(x, y, z, m, v, train_ticks, samples, dt, theta, leaf_cap, seed=seed, log=log)
    vs = val_system if val_system is not None else (x, y, z, m, v)
    va = collect(*vs, val_ticks, samples, dt, theta, leaf_cap, seed=seed + 7777, log=log)
    params, trainer, hist = train(tr, steps, batch, lr, seed, log)
    pred = trainer.forward(va["F"])
    trainer.close()
    e_plain = force_rms(va)
    e_corr = force_rms(va, pred)
    _log(log, f"mlp: validation FP64 RMS plain {e_plain:.5e}, corrected {e_corr:.5e}")
    return dict(params=params, val_plain=e_plain, val_corrected=e_corr,
                enabled=bool(e_corr < e_plain), train_samples=int(tr["F"].shape[0]),
                val_samples=int(va["F"].shape[0]), final_loss=float(np.mean(hist[-100:])),
                train_plain=force_rms(tr))


def stderr_log(msg):
    print(msg, file=sys.stderr, flush=True)
