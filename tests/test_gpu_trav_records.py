"""The traversal's two ways of writing a next list: from the near records
the count phase keeps in LDS (the default), and from the list, far mask and
links again (what an entry whose records did not fit in its group's share
of the pool gets; PS_HIP_OPT_TRAV_RECORDS = 0 forces it for every entry).
Bar: the tick is the same bit for bit either way (sorted state, forces,
tree, pair and list counts), and both agree with the oracle's pair counts."""
import numpy as np
import pytest

from gpu_helpers import log
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("codes", "id", "x", "y", "z", "m", "fx", "fy", "fz")


def tick(x, y, z, m, records, trav=api.TRAV_MT, lanes=None, leaf_cap=None, kahan=False):
    with api.Solver(len(x), theta=2.0, traversal=trav, stats=True, kahan=kahan) as s:
        api.check("opt", api.ps_set_option(s.ctx, api.OPT_TRAV_RECORDS, records))
        if lanes is not None:
            api.check("lanes", api.ps_set_option(s.ctx, api.OPT_TRAV_LANES, lanes))
        if leaf_cap is not None:
            api.check("cap", api.ps_set_option(s.ctx, api.OPT_LEAF_CAP, leaf_cap))
        s.load(x, y, z, m)
        s.tick(api.MORTON_AVX2)
        return s.result(), s.stats(), s.tree()


@pytest.mark.parametrize("case", ["plummer", "two-plummers-st", "cube", "lanes16", "cap8",
                                  "kahan"])
def test_records_and_lists_write_the_same_tick(case):
    n = 150_000
    kw = {}
    if case == "two-plummers-st":
        x, y, z, m = W.two_plummers(n, seed=601)[:4]
        kw["trav"] = api.TRAV_ST
    elif case == "cube":
        x, y, z, m = W.uniform_cube(n, seed=603, mass=1.0 / n)
    else:
        x, y, z, m = W.plummer(n, seed=602)
    if case == "lanes16":
        kw["lanes"] = 16
    if case == "cap8":
        kw["leaf_cap"] = 8
    if case == "kahan":
        kw["kahan"] = True
    a, sa, ta = tick(x, y, z, m, 1, **kw)
    b, sb, tb = tick(x, y, z, m, 0, **kw)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    for k in ("n_nodes", "m2l_pairs", "p2p_pairs", "p2p_leaf_pairs", "list_entries"):
        assert sa[k] == sb[k], k
    for k in ("geom", "mp", "loc", "link", "nend", "parent"):
        assert np.array_equal(ta[k], tb[k]), k
    log(f"{case}: records = lists, m2l {sa['m2l_pairs']} list words {sa['list_entries']}")


def test_option_is_checked():
    with api.Solver(1000, theta=2.0) as s:
        assert api.ps_set_option(s.ctx, api.OPT_TRAV_RECORDS, 2) == api.PS_EINVAL
