"""The CPU check of the learned correction, pinned by mathematics: its
backward pass against central finite differences of its forward pass, GELU
(tanh form) against its definition and the erf form, and the packing round
trip."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_ref as R  # noqa: E402


def test_param_count_and_packing():
    assert R.N_PARAMS == 10115
    p = np.arange(R.N_PARAMS, dtype=np.float64)
    assert np.array_equal(R.pack(R.unpack(p)), p)


def test_gelu_tanh_form():
    c = math.sqrt(2 / math.pi)
    for x in (-30.0, -3.0, -0.5, 0.0, 0.7, 2.5, 30.0):
        want = 0.5 * x * (1 + math.tanh(c * (x + 0.044715 * x ** 3)))
        assert abs(R.gelu(np.array(x)) - want) <= 1e-15 * max(1.0, abs(x))
        # the tanh form stays within 1e-3 of the exact erf form
        assert abs(R.gelu(np.array(x)) - 0.5 * x * (1 + math.erf(x / math.sqrt(2)))) < 1e-3
        h = 1e-6
        fd = (R.gelu(np.array(x + h)) - R.gelu(np.array(x - h))) / (2 * h)
        assert abs(R.gelu_d(np.array(x)) - fd) < 1e-8
    # limits: 0 far left, identity far right
    assert R.gelu(np.array(-30.0)) == 0.0 and R.gelu(np.array(30.0)) == 30.0


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = R.init_params(seed=5, out_scale=1.0).astype(np.float64)
    f = rng.normal(size=(7, R.N_IN))
    t = rng.normal(size=(7, R.N_OUT))
    loss, g = R.loss_grad(params, f, t)
    idx = rng.choice(R.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [R.N_PARAMS - 1, R.N_PARAMS - 4, 0, 1535, 1536]])
    for i in idx:
        h = 1e-6
        p1 = params.copy()
        p1[i] += h
        p2 = params.copy()
        p2[i] -= h
        fd = (R.loss_grad(p1, f, t)[0] - R.loss_grad(p2, f, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
