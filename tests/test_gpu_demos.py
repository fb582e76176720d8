"""The reference's raylib demos (demo/unary.c, demo/binary.c) as headless C
programs against the drop-in, including their call into the reference's
internals, ps_impl_arena_clear(&ctx->arena) (demo/unary.c:120).  The first
frame must match the oracle on the same inputs (codes, ids and payload bit
for bit, forces within the parity tolerance) and the demo's 60 frames must
stay finite with its black holes in place."""
import os
import subprocess

import numpy as np
import pytest

from gpu_helpers import FORCE_RMS_TOL, forces, log, rms_rel
from polesitter_amd import api

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n,trav,holes", [("demo_unary", 40000, 0, 1),
                                               ("demo_binary", 10000, 1, 2)])
def test_headless_demo(oracle, tmp_path, name, n, trav, holes):
    out = tmp_path / f"{name}.bin"
    r = subprocess.run([os.path.join(api.LIBDIR, name), str(out), "60"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    log(r.stdout.strip())
    summary = dict(kv.split("=") for kv in r.stdout.split()[1:])
    assert summary["finite"] == "1" and summary["black_holes"] == str(holes)
    raw = np.fromfile(out, dtype=np.uint32)
    assert raw.size == 13 * n
    f = raw.view(np.float32)
    x0, y0, z0, m0 = (f[k * n:(k + 1) * n] for k in range(4))
    got = {k: f[(4 + j) * n:(5 + j) * n] for j, k in enumerate(("x", "y", "z", "m", "fx", "fy",
                                                               "fz"))}
    got["id"] = raw[11 * n:12 * n]
    got["codes"] = raw[12 * n:13 * n]
    o = oracle.tick(x0, y0, z0, m0, theta=2.0, trav=trav)
    for k in ("codes", "id", "x", "y", "z", "m"):
        assert np.array_equal(got[k], o[k]), k
    err = rms_rel(forces(got), forces(o))
    log(f"{name}: first-frame force RMS vs oracle {err:.2e}")
    assert err <= FORCE_RMS_TOL
