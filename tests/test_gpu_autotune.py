"""The autotuner on a small mixed ensemble: the reference setting is among
the measured candidates, the chosen setting is no slower than it and no
less accurate in any size class (by the tuner's own measurements), and a
direct run of the chosen setting equals, system by system, single ticks
with each system's chosen leaf capacity."""
import numpy as np
import pytest

from gpu_helpers import gpu_tick, log
from polesitter_amd import api, autotune
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_autotuner_per_class_choice_is_honest():
    systems = W.ensemble(20, seed=9, lo=2048, hi=60000)
    sizes = [len(s[0]) for s in systems]
    classes = autotune.size_classes(sizes)
    assert len(set(classes.tolist())) >= 2
    r = autotune.tune_batch(systems, reps=2, error_systems_per_class=3,
                            targets_per_system=128, log=log)
    ref = r["reference"]
    cfgs = [(tuple(c["caps"]), c["p2p_big"], c["trav_lanes"]) for c, *_ in r["table"]]
    assert (tuple(ref["config"]["caps"]), 24, 8) in cfgs
    assert r["ms_tick"] <= ref["ms_tick"]
    for c, e in r["class_errors"].items():
        assert e <= ref["class_errors"][c] * (1 + 1e-6), (c, e, ref["class_errors"][c])
    assert len(r["system_caps"]) == len(systems)
    log(f"autotune: chosen {r['config']} {r['ms_tick']:.3f} ms vs reference "
        f"{ref['ms_tick']:.3f} ms")
    # the chosen setting run directly: every system the single tick of its
    # own capacity
    n = sum(sizes)
    err_sys = [k for k in range(len(systems))][:: max(1, len(systems) // 6)]
    with api.Batch(n, len(systems), theta=2.0, graph=False) as b:
        b.set_leaf_caps(r["system_caps"])
        b.set_option(api.OPT_P2P_BIG, r["config"]["p2p_big"])
        b.set_option(api.OPT_TRAV_LANES, r["config"]["trav_lanes"])
        b.load(systems)
        b.tick(api.MORTON_AVX2)
        res = b.result()
    for k in err_sys[:3]:
        x, y, z, m = systems[k]
        o = gpu_tick(x, y, z, m, theta=2.0, mode=api.MORTON_AVX2, tree=False,
                     leaf_cap=r["system_caps"][k], lanes=r["config"]["trav_lanes"],
                     p2p_big=r["config"]["p2p_big"])
        g = b.system(res, k)
        for f in ("codes", "id", "fx", "fy", "fz"):
            assert np.array_equal(g[f], o[f]), (k, f)
