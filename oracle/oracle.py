"""ctypes front end of the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product library never routes
through it.

Two kinds of CPU code are reachable from here:

* ``lib/libps_oracle.so`` -- this repository's C restatement of the
  reference tick (ps_oracle.c), bit-identical to a strict build of
  src/polesitter.h.
* ``_ref/libref_*.so`` -- the reference's own header compiled by
  oracle/Makefile (present where the reference tree was available at build
  time; the files travel with the repository snapshot).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "lib", "libps_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

TRAV_MT = 0  # jobs (target at depth 2 or shallower leaf, root), H:1772-1801
TRAV_ST = 1  # one pair (root, root), H:1888
MORTON_TRUNC = 0  # scalar path H:2132-2157
MORTON_AVX2 = 1  # 8-wide path H:2040-2072 + scalar tail

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class _Nodes(C.Structure):
    _fields_ = [
        ("cap", C.c_uint32), ("n_nodes", C.c_uint32), ("max_depth", C.c_uint32),
        ("level_start", C.c_uint32 * 12),
        ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("hw", C.c_void_p),
        ("first", C.c_void_p), ("end", C.c_void_p), ("child", C.c_void_p),
        ("mask", C.c_void_p), ("parent", C.c_void_p), ("depth", C.c_void_p),
        ("mp", C.c_void_p), ("loc", C.c_void_p),
    ]


class _Counts(C.Structure):
    _fields_ = [("m2l_pairs", C.c_uint64), ("p2p_leaf_pairs", C.c_uint64),
                ("p2p_pairs", C.c_uint64), ("visited_pairs", C.c_uint64)]


class Tree:
    """Canonical level-order node arrays (numpy), filled by the oracle."""

    def __init__(self, cap: int):
        self.cap = int(cap)
        f = lambda k=1: np.zeros(self.cap * k, np.float32)
        u = lambda: np.zeros(self.cap, np.uint32)
        self.x, self.y, self.z, self.hw = f(), f(), f(), f()
        self.first, self.end, self.child, self.mask, self.parent = u(), u(), u(), u(), u()
        self.depth = np.zeros(self.cap, np.uint8)
        self.mp, self.loc = f(4), f(4)
        self.st = _Nodes()
        self.st.cap = self.cap
        for name in ("x", "y", "z", "hw", "first", "end", "child", "mask",
                     "parent", "depth", "mp", "loc"):
            setattr(self.st, name, getattr(self, name).ctypes.data)

    @property
    def n(self) -> int:
        return int(self.st.n_nodes)

    @property
    def level_start(self) -> np.ndarray:
        return np.array(list(self.st.level_start), np.uint32)

    def view(self) -> dict:
        n = self.n
        return dict(
            n_nodes=n, level_start=self.level_start,
            geom=np.stack([self.x[:n], self.y[:n], self.z[:n], self.hw[:n]], 1),
            first=self.first[:n].copy(), end=self.end[:n].copy(),
            child=self.child[:n].copy(), mask=self.mask[:n].copy(),
            parent=self.parent[:n].copy(), depth=self.depth[:n].copy(),
            leaf=(self.mask[:n] == 0), mp=self.mp[: 4 * n].reshape(n, 4).copy(),
            loc=self.loc[: 4 * n].reshape(n, 4).copy(),
        )


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"oracle library missing: {ORACLE_LIB} (run make -C oracle)")
        L = C.CDLL(ORACLE_LIB)
        L.or_prepare.argtypes = [C.c_size_t, _f32p, _f32p, _f32p, _u32p, C.c_int,
                                 C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.or_prepare.restype = C.c_int
        L.or_expand_bits.argtypes = [C.c_uint32]
        L.or_expand_bits.restype = C.c_uint32
        L.or_morton_encode.argtypes = [C.c_uint32] * 3
        L.or_morton_encode.restype = C.c_uint32
        L.or_sort.argtypes = [C.c_size_t, _u32p, _u32p, _f32p, _f32p, _f32p, _f32p]
        L.or_build.argtypes = [C.c_size_t, _u32p, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.POINTER(_Nodes)]
        L.or_build.restype = C.c_long
        L.or_upward.argtypes = [C.POINTER(_Nodes), _f32p, _f32p, _f32p, _f32p]
        L.or_interaction.argtypes = [C.POINTER(_Nodes), C.c_float, C.c_int, C.POINTER(_Counts)]
        L.or_downward.argtypes = [C.POINTER(_Nodes), _f32p, _f32p, _f32p, _f32p]
        L.or_p2p.argtypes = [C.POINTER(_Nodes), C.c_float, C.c_int, _f32p, _f32p, _f32p, _f32p,
                             _f32p, _f32p, _f32p, C.POINTER(_Counts)]
        L.or_calc_forces.argtypes = [C.c_size_t, _f32p, _f32p, _f32p, _f32p, _u32p, _u32p,
                                     _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_int, C.POINTER(_Nodes),
                                     C.POINTER(_Counts)]
        L.or_calc_forces.restype = C.c_int
        L.or_direct_f64.argtypes = [C.c_size_t, _f32p, _f32p, _f32p, _f32p, C.c_size_t,
                                    C.c_void_p, _f64p, _f64p, _f64p]
        L.or_mac_far.argtypes = [C.c_float] * 9
        L.or_mac_far.restype = C.c_int
        L.or_m2l.argtypes = [C.c_float] * 6 + [_f32p, _f32p]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def prepare(x, y, z, mode=MORTON_TRUNC):
    """ps_prepare_particles restated: returns (codes, min, max, range)."""
    x, y, z = _f32(x), _f32(y), _f32(z)
    codes = np.zeros(x.size, np.uint32)
    lo, hi, rg = C.c_float(), C.c_float(), C.c_float()
    rc = lib().or_prepare(x.size, x, y, z, codes, mode, C.byref(lo), C.byref(hi), C.byref(rg))
    if rc != 0:
        raise ValueError(f"or_prepare returned {rc}")
    return codes, lo.value, hi.value, rg.value


def morton_encode(x, y, z) -> int:
    return int(lib().or_morton_encode(x, y, z))


def node_cap(n: int) -> int:
    """Upper bound on the node count for n particles (one start per depth)."""
    return int(1 + 10 * max(int(n), 1))


def calc_forces(x, y, z, m, ids, codes, root_c, root_hw, theta=2.0, trav=TRAV_MT,
                cap=None):
    """ps_calc_forces restated, on copies.  Returns a dict with the sorted
    arrays, forces, the canonical tree and traversal counters."""
    x, y, z, m = _f32(x).copy(), _f32(y).copy(), _f32(z).copy(), _f32(m).copy()
    ids = np.ascontiguousarray(ids, np.uint32).copy()
    codes = np.ascontiguousarray(codes, np.uint32).copy()
    n = x.size
    fx, fy, fz = (np.zeros(n, np.float32) for _ in range(3))
    t = Tree(cap or node_cap(n))
    cnt = _Counts()
    cx, cy, cz = (root_c if np.ndim(root_c) else (root_c,) * 3)
    rc = lib().or_calc_forces(n, x, y, z, m, ids, codes, fx, fy, fz, cx, cy, cz, root_hw,
                              theta, trav, C.byref(t.st), C.byref(cnt))
    if rc != 0:
        raise RuntimeError(f"or_calc_forces returned {rc}")
    return dict(x=x, y=y, z=z, m=m, id=ids, codes=codes, fx=fx, fy=fy, fz=fz,
                tree=t.view(), counts=dict(m2l_pairs=cnt.m2l_pairs,
                                           p2p_leaf_pairs=cnt.p2p_leaf_pairs,
                                           p2p_pairs=cnt.p2p_pairs,
                                           visited_pairs=cnt.visited_pairs))


def tick(x, y, z, m, theta=2.0, trav=TRAV_MT, mode=MORTON_TRUNC, ids=None):
    """prepare + calc_forces as the reference's callers do (accuracy.c:493-503)."""
    codes, lo, hi, rg = prepare(x, y, z, mode)
    if ids is None:
        ids = np.arange(len(x), dtype=np.uint32)
    root_c = np.float32(np.float32(lo) + np.float32(rg) * np.float32(0.5))
    out = calc_forces(x, y, z, m, ids, codes, float(root_c), float(np.float32(rg) * np.float32(0.5)),
                      theta, trav)
    out.update(prep_codes=codes, min=lo, max=hi, range=rg)
    return out


def direct_f64(x, y, z, m, targets=None):
    """FP64 direct sum of tests/accuracy.c:42-72 for the given target indices."""
    x, y, z, m = _f32(x), _f32(y), _f32(z), _f32(m)
    if targets is None:
        tg = None
        nt = x.size
    else:
        tg = np.ascontiguousarray(targets, np.uint32)
        nt = tg.size
    fx, fy, fz = (np.zeros(nt, np.float64) for _ in range(3))
    lib().or_direct_f64(x.size, x, y, z, m, nt, None if tg is None else tg.ctypes.data,
                        fx, fy, fz)
    return np.stack([fx, fy, fz], 1)


def error_metrics(f32_forces: np.ndarray, ref64: np.ndarray):
    """RMS and max relative error, tests/accuracy.c:505-531."""
    d = f32_forces.astype(np.float64) - ref64
    e2 = (d * d).sum(1)
    r2 = (ref64 * ref64).sum(1)
    rms = float(np.sqrt(e2.sum() / r2.sum())) if r2.sum() > 0 else 0.0
    ok = r2 > 1e-24
    mx = float(np.sqrt(e2[ok] / r2[ok]).max()) if ok.any() else 0.0
    return rms, mx


# ---------------------------------------------------------------------------
# the compiled reference (oracle/_ref), when present
# ---------------------------------------------------------------------------

class _RefDump(C.Structure):
    _fields_ = [("cap", C.c_uint32), ("n_nodes", C.c_uint32),
                ("depth", C.c_void_p), ("leaf", C.c_void_p), ("first", C.c_void_p),
                ("cnt", C.c_void_p), ("parent", C.c_void_p), ("geom", C.c_void_p),
                ("mp", C.c_void_p), ("loc", C.c_void_p)]


REF_BUILDS = ("strict_mt", "strict_st", "shipped_mt")


def ref_available(build="strict_mt") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libref_{build}.so"))


_ref_libs = {}


def ref_lib(build="strict_mt"):
    if build not in _ref_libs:
        path = os.path.join(REF_DIR, f"libref_{build}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference build missing: {path}")
        L = C.CDLL(path)
        L.ref_prepare.argtypes = [C.c_size_t, _f32p, _f32p, _f32p, _u32p, _f32p]
        L.ref_prepare.restype = C.c_int
        args = [C.c_size_t, _f32p, _f32p, _f32p, _f32p, _u32p, _u32p, _f32p, _f32p, _f32p]
        L.ref_tick.argtypes = args + [C.c_float, C.c_size_t, C.c_size_t, _f32p,
                                      C.POINTER(_RefDump)]
        L.ref_tick.restype = C.c_int
        L.ref_calc.argtypes = args + [C.c_float] * 5 + [C.c_size_t, C.c_size_t,
                                                        C.POINTER(_RefDump)]
        L.ref_calc.restype = C.c_int
        L.ref_time_ticks.argtypes = args + [C.c_float, C.c_size_t, C.c_size_t, C.c_int, _f64p]
        L.ref_time_ticks.restype = C.c_int
        _ref_libs[build] = L
    return _ref_libs[build]


def ref_buff_size(n: int) -> int:
    """A buffer the reference accepts for n particles: context + radix temp
    (24 B/particle, H:1717) + room for ~2.5 nodes of 80 B per particle."""
    return int(4 << 20) + 24 * n + 200 * max(n, 1024)


def ref_tick(x, y, z, m, theta=2.0, thrd_cnt=4, build="strict_mt", dump=True, ids=None,
             buff_size=None):
    """Run the genuine reference tick (prepare + calc_forces) on copies."""
    L = ref_lib(build)
    x, y, z, m = _f32(x).copy(), _f32(y).copy(), _f32(z).copy(), _f32(m).copy()
    n = x.size
    ids = (np.arange(n, dtype=np.uint32) if ids is None
           else np.ascontiguousarray(ids, np.uint32).copy())
    codes = np.zeros(n, np.uint32)
    fx, fy, fz = (np.zeros(n, np.float32) for _ in range(3))
    out3 = np.zeros(3, np.float32)
    d = None
    keep = {}
    if dump:
        cap = node_cap(n)
        keep = dict(depth=np.zeros(cap, np.uint8), leaf=np.zeros(cap, np.uint8),
                    first=np.zeros(cap, np.uint32), cnt=np.zeros(cap, np.uint32),
                    parent=np.zeros(cap, np.uint32), geom=np.zeros(4 * cap, np.float32),
                    mp=np.zeros(4 * cap, np.float32), loc=np.zeros(4 * cap, np.float32))
        d = _RefDump()
        d.cap = cap
        for k, v in keep.items():
            setattr(d, k, v.ctypes.data)
    rc = L.ref_tick(n, x, y, z, m, ids, codes, fx, fy, fz, theta, thrd_cnt,
                    buff_size or ref_buff_size(n), out3, C.byref(d) if d is not None else None)
    if rc < 0:
        raise RuntimeError(f"ref_tick returned {rc}")
    out = dict(x=x, y=y, z=z, m=m, id=ids, codes=codes, fx=fx, fy=fy, fz=fz,
               min=float(out3[0]), max=float(out3[1]), range=float(out3[2]))
    if dump:
        nn = int(d.n_nodes)
        out["tree"] = dict(
            n_nodes=nn, depth=keep["depth"][:nn], leaf=keep["leaf"][:nn].astype(bool),
            first=keep["first"][:nn], cnt=keep["cnt"][:nn], parent=keep["parent"][:nn],
            geom=keep["geom"][: 4 * nn].reshape(nn, 4), mp=keep["mp"][: 4 * nn].reshape(nn, 4),
            loc=keep["loc"][: 4 * nn].reshape(nn, 4))
    return out


def ref_prepare(x, y, z, build="strict_mt"):
    x, y, z = _f32(x).copy(), _f32(y).copy(), _f32(z).copy()
    codes = np.zeros(x.size, np.uint32)
    out3 = np.zeros(3, np.float32)
    rc = ref_lib(build).ref_prepare(x.size, x, y, z, codes, out3)
    if rc != 0:
        raise RuntimeError(f"ref_prepare returned {rc}")
    return codes, float(out3[0]), float(out3[1]), float(out3[2])


def ref_time_ticks(x, y, z, m, theta=2.0, thrd_cnt=4, n_ticks=3, build="shipped_mt",
                   buff_size=None):
    """Wall seconds per tick of the reference CPU path (benchmark.c:175-184)."""
    L = ref_lib(build)
    x, y, z, m = _f32(x).copy(), _f32(y).copy(), _f32(z).copy(), _f32(m).copy()
    n = x.size
    ids = np.arange(n, dtype=np.uint32)
    codes = np.zeros(n, np.uint32)
    fx, fy, fz = (np.zeros(n, np.float32) for _ in range(3))
    secs = np.zeros(n_ticks, np.float64)
    rc = L.ref_time_ticks(n, x, y, z, m, ids, codes, fx, fy, fz, theta, thrd_cnt,
                          buff_size or ref_buff_size(n), n_ticks, secs)
    if rc != 0:
        raise RuntimeError(f"ref_time_ticks returned {rc}")
    return secs
