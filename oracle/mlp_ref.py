"""CPU check of the learned far-field correction (test infrastructure only:
tests/ use it to check the HIP kernels; the product path never calls it).

The correction has no counterpart in the reference; what pins this file is
mathematics: the backward pass is checked against central finite
differences of the forward pass (tests/test_oracle_mlp.py), and GELU (tanh
form) against its definition and the erf form it approximates.  Everything
is float64.

Layout of the packed parameters (same as the kernels):
    W1[64][24] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 24, 64, 3
SHAPES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H, N_H)), ("b3", (N_H,)), ("W4", (N_OUT, N_H)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
NONE = 0xFFFFFFFF


def unpack(params):
    p = np.asarray(params, np.float64)
    assert p.size == N_PARAMS
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def gelu(x):
    """GELU, tanh form: x/2 (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))"""
    return 0.5 * x * (1.0 + np.tanh(GELU_C * (x + GELU_A * x ** 3)))


def gelu_d(x):
    t = np.tanh(GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def forward(params, feat):
    """net(feat) for feat[n][24]; returns (out[n][3], cache)"""
    w = unpack(params)
    f = np.asarray(feat, np.float64)
    z1 = f @ w["W1"].T + w["b1"]
    h1 = gelu(z1)
    z2 = h1 @ w["W2"].T + w["b2"]
    h2 = gelu(z2)
    z3 = h2 @ w["W3"].T + w["b3"]
    h3 = gelu(z3)
    out = h3 @ w["W4"].T + w["b4"]
    return out, dict(f=f, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, w=w)


def loss_grad(params, feat, target):
    """loss = mean over samples of |net(F) - t|^2 and its gradient (packed)"""
    out, c = forward(params, feat)
    n = out.shape[0]
    d = out - np.asarray(target, np.float64)
    loss = float((d * d).sum() / n)
    w = c["w"]
    d4 = 2.0 * d / n
    g = {"W4": d4.T @ c["h3"], "b4": d4.sum(0)}
    d3 = (d4 @ w["W4"]) * gelu_d(c["z3"])
    g["W3"], g["b3"] = d3.T @ c["h2"], d3.sum(0)
    d2 = (d3 @ w["W3"]) * gelu_d(c["z2"])
    g["W2"], g["b2"] = d2.T @ c["h1"], d2.sum(0)
    d1 = (d2 @ w["W2"]) * gelu_d(c["z1"])
    g["W1"], g["b1"] = d1.T @ c["f"], d1.sum(0)
    return loss, pack(g)


def init_params(seed=0, out_scale=0.1):
    """He-style initialisation; a small last layer so the untrained
    correction starts near zero"""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("b"):
            d[name] = np.zeros(shape)
        else:
            fan_in = shape[1]
            d[name] = rng.normal(0.0, np.sqrt(2.0 / fan_in), shape)
    d["W4"] *= out_scale
    return pack(d).astype(np.float32)


def features(pos4, leaf_of, geom, mp, loc, parent):
    """the feature vector and scale of every particle (sorted order), as
    ps_mlp.hip's mlp_features computes them (float64 here)"""
    p = np.asarray(pos4, np.float64).reshape(-1, 4)
    L = np.asarray(leaf_of, np.int64)
    par = np.asarray(parent, np.uint32)
    geom = np.asarray(geom, np.float64).reshape(-1, 4)
    mp = np.asarray(mp, np.float64).reshape(-1, 4)
    loc = np.asarray(loc, np.float64).reshape(-1, 4)
    n = p.shape[0]
    P = par[L].astype(np.int64)
    hasP = par[L] != NONE
    Pi = np.where(hasP, P, L)
    G = np.where(hasP, par[Pi], NONE)
    hasG = G != NONE
    Gi = np.where(hasG, G.astype(np.int64), Pi)
    E = loc[L, 1:]
    EP = np.where(hasP[:, None], loc[Pi, 1:], 0.0)
    EG = np.where(hasG[:, None], loc[Gi, 1:], 0.0)
    s = np.sqrt((E * E).sum(1))
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    f = np.zeros((n, N_IN))
    f[:, 0:3] = (p[:, :3] - geom[L, :3]) / geom[L, 3:4]
    f[:, 3:6] = (E - EP) * inv[:, None]
    f[:, 6:9] = np.where(hasP[:, None], (p[:, :3] - geom[Pi, :3]) / geom[Pi, 3:4], 0.0)
    f[:, 9:12] = (EP - EG) * inv[:, None]
    f[:, 12:15] = np.where(hasG[:, None], (p[:, :3] - geom[Gi, :3]) / geom[Gi, 3:4], 0.0)
    f[:, 15:18] = EG * inv[:, None]
    m0 = mp[L, 0]
    cm = np.where(m0 != 0, 1.0 / (geom[L, 3] * np.where(m0 != 0, m0, 1.0)), 0.0)
    f[:, 21:24] = mp[L, 1:] * cm[:, None]
    f[s == 0] = 0.0
    return f, s
