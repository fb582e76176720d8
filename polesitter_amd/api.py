"""Python mirror of the polesitter C API, bound with ctypes.

Names, argument meaning and result codes follow the reference header
(src/polesitter.h:294-339): ``ps_init``, ``ps_prepare_particles``,
``ps_calc_forces``, ``ps_destroy``, ``PsConfig`` / ``PsParticleArrs`` with the
reference's field order.  The functions call libpolesitter.so (the C host
layer of include/polesitter.h), which drives the gfx950 kernels in
libpolesitter_hip.so.  There is no CPU fallback: importing this module
without the built libraries raises.

Device memory for callers that keep particles on the GPU comes from
``DeviceArray`` (ps_hip_malloc); numpy arrays are host memory and take the
library's staging path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HIP_LIB = os.path.join(LIBDIR, "libpolesitter_hip.so")
HOST_LIB = os.path.join(LIBDIR, "libpolesitter.so")

PS_OK, PS_EOOM, PS_EINVAL, PS_ETHRD, PS_EALLOC, PS_EDEVICE = 0, -1, -2, -3, -4, -5
RESULT_NAMES = {0: "PS_OK", -1: "PS_EOOM", -2: "PS_EINVAL", -3: "PS_ETHRD", -4: "PS_EALLOC",
                -5: "PS_EDEVICE"}

TRAV_MT, TRAV_ST = 0, 1
MORTON_TRUNC, MORTON_AVX2 = 0, 1
OPT_TRAVERSAL, OPT_KAHAN, OPT_THETA, OPT_STATS, OPT_TRAV_LANES, OPT_P2P_BIG, OPT_LEAF_CAP = \
    1, 2, 3, 4, 5, 6, 7
OPT_ARENA_WORDS = 8
OPT_TRAV_BLOCKS = 9
OPT_TRAV_SEGMENTS = 10
OPT_SORT_INCREMENTAL = 11
OPT_TRAV_RECORDS = 12
BATCH_OPT_GRAPH = 100


class PsError(RuntimeError):
    def __init__(self, where, rc):
        super().__init__(f"{where} returned {RESULT_NAMES.get(rc, rc)}")
        self.rc = rc


class PsConfig(C.Structure):
    """ps_config_t, src/polesitter.h:304-310"""
    _fields_ = [("buff", C.c_void_p), ("buff_size", C.c_size_t),
                ("max_particles", C.c_size_t), ("theta", C.c_float), ("thrd_cnt", C.c_size_t)]


class PsConfigEx(C.Structure):
    """ps_config_ex_t: ps_config_t followed by caller-owned device memory"""
    _fields_ = [("base", PsConfig), ("dev_buff", C.c_void_p), ("dev_buff_size", C.c_size_t)]


class PsParticleArrs(C.Structure):
    """ps_particle_arrs_t, src/polesitter.h:312-322"""
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("mass", C.c_void_p), ("fx", C.c_void_p), ("fy", C.c_void_p),
                ("fz", C.c_void_p), ("id", C.c_void_p), ("cnt", C.c_size_t)]


class PsBatchStats(C.Structure):
    _fields_ = [("n_particles", C.c_uint64), ("n_systems", C.c_uint64), ("n_nodes", C.c_uint64),
                ("m2l_pairs", C.c_uint64), ("graph_replays", C.c_uint64),
                ("used_graph", C.c_uint32), ("ms_tick", C.c_float)]


class PsStats(C.Structure):
    """ps_hip_stats_t, include/polesitter_hip.h"""
    _fields_ = [("n_particles", C.c_uint64), ("n_nodes", C.c_uint64), ("n_leaves", C.c_uint64),
                ("max_depth", C.c_uint32), ("iterations", C.c_uint32),
                ("level_start", C.c_uint32 * 12),
                ("m2l_pairs", C.c_uint64), ("p2p_leaf_pairs", C.c_uint64),
                ("p2p_pairs", C.c_uint64), ("list_entries", C.c_uint64),
                ("ms_upload", C.c_float), ("ms_sort", C.c_float), ("ms_build", C.c_float),
                ("ms_upward", C.c_float), ("ms_traverse", C.c_float), ("ms_l2l", C.c_float),
                ("ms_p2p", C.c_float), ("ms_download", C.c_float), ("ms_total", C.c_float),
                ("ms_trav_steps", C.c_float), ("ms_mlp", C.c_float),
                ("sort_changed", C.c_uint64), ("sort_incremental", C.c_uint32),
                ("reserved_", C.c_uint32)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name == "level_start" else v
        return d


_hip = None
_host = None

_vp = C.c_void_p
_sz = C.c_size_t
_f = C.c_float


def hip_lib():
    """libpolesitter_hip.so with its ps_hip_* signatures."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError(f"{HIP_LIB} is not built (python -m polesitter_amd.build)")
        L = C.CDLL(HIP_LIB, mode=C.RTLD_GLOBAL)
        sig = {
            "ps_hip_engine_create": ([C.POINTER(_vp), _sz, _sz, _f, C.c_int], C.c_int),
            "ps_hip_engine_destroy": ([_vp], C.c_int),
            "ps_hip_engine_set_option": ([_vp, C.c_int, C.c_double], C.c_int),
            "ps_hip_engine_get_stats": ([_vp, C.POINTER(PsStats)], C.c_int),
            "ps_hip_prepare": ([_vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp, C.c_int], C.c_int),
            "ps_hip_prepare_noalloc": ([_vp, _vp, _vp, _sz, _vp, _vp, _vp, _vp, C.c_int],
                                       C.c_int),
            "ps_hip_calc_forces": ([_vp] * 9 + [_sz, _vp, _f, _f, _f, _f], C.c_int),
            "ps_hip_engine_export_tree": ([_vp, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp], C.c_int),
            "ps_hip_debug_bind": ([_vp] * 9 + [_sz, _vp, _f, _f, _f, _f], C.c_int),
            "ps_hip_debug_stage": ([_vp, C.c_int], C.c_int),
            "ps_hip_debug_set_local": ([_vp, C.c_uint32, _vp], C.c_int),
            "ps_hip_op_m2l": ([_vp, _vp, _vp, _f, _vp], C.c_int),
            "ps_hip_op_morton_exhaustive": ([_vp], C.c_int),
            "ps_hip_op_copy": ([_vp, _vp, C.c_size_t], C.c_int),
            "ps_hip_op_bandwidth": ([C.c_int, _vp, _vp, _vp, C.c_size_t], C.c_int),
            "ps_hip_engine_set_mlp": ([_vp, _vp], C.c_int),
            "ps_hip_mlp_features": ([_vp, _vp, _vp], C.c_int),
            "ps_hip_mlp_features_at": ([_vp, _vp, _sz, _vp, _vp], C.c_int),
            "ps_hip_mlp_forward": ([_vp, _vp, _vp, _sz, _vp], C.c_int),
            "ps_hip_mlp_grad": ([_vp, _vp, _vp, _vp, _sz, _vp, C.POINTER(_f)], C.c_int),
            "ps_hip_mlp_adam": ([_vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, C.c_int], C.c_int),
            "ps_hip_batch_create": ([C.POINTER(_vp), _sz, _sz, _f, C.c_int], C.c_int),
            "ps_hip_batch_destroy": ([_vp], C.c_int),
            "ps_hip_batch_set_option": ([_vp, C.c_int, C.c_double], C.c_int),
            "ps_hip_batch_tick": ([_vp] * 11 + [_sz, C.c_int], C.c_int),
            "ps_hip_batch_get_stats": ([_vp, C.POINTER(PsBatchStats)], C.c_int),
            "ps_hip_batch_set_leaf_caps": ([_vp, _vp, _sz], C.c_int),
            "ps_hip_direct_f64": ([_vp, _vp, _vp, _vp, _sz, _vp, _sz, _vp], C.c_int),
            "ps_hip_step_reset_ids": ([_vp, _vp, _sz], C.c_int),
            "ps_hip_step_gather": ([_vp, _vp, _vp, _vp, _vp, _vp, _sz], C.c_int),
            "ps_hip_step_euler": ([_vp] * 11 + [_f, _sz], C.c_int),
            "ps_hip_step_frame": ([_vp] * 13 + [_f, _sz], C.c_int),
            "ps_hip_malloc": ([_sz], _vp),
            "ps_hip_free": ([_vp], C.c_int),
            "ps_hip_memcpy": ([_vp, _vp, _sz, C.c_int], C.c_int),
            "ps_hip_synchronize": ([], C.c_int),
            "ps_hip_stream_create": ([C.POINTER(_vp)], C.c_int),
            "ps_hip_stream_destroy": ([_vp], C.c_int),
            "ps_hip_stream_synchronize": ([_vp], C.c_int),
            "ps_hip_engine_create_in": ([C.POINTER(_vp), _sz, _sz, _f, C.c_int, _vp, _sz, _vp,
                                         _sz], C.c_int),
            "ps_hip_engine_host_bytes": ([], _sz),
            "ps_hip_engine_scratch_bytes": ([_sz], _sz),
            "ps_hip_worst_case_nodes": ([_sz], _sz),
            "ps_hip_alloc_stats": ([C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
            "ps_hip_debug_fail": ([C.c_int, C.c_int], C.c_int),
            "ps_hip_calc_forces_async": ([_vp] * 9 + [_sz, _vp, _f, _f, _f, _f, _vp], C.c_int),
            "ps_hip_calc_forces_wait": ([_vp], C.c_int),
            "ps_hip_device_name": ([C.c_char_p, _sz], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _hip = L
    return _hip


def host_lib():
    """libpolesitter.so: ps_init / ps_prepare_particles / ps_calc_forces / ps_destroy."""
    global _host
    if _host is None:
        hip_lib()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is not built (python -m polesitter_amd.build)")
        L = C.CDLL(HOST_LIB)
        L.ps_init.argtypes = [C.POINTER(_vp), C.POINTER(PsConfig)]
        L.ps_init.restype = C.c_int
        L.ps_calc_forces.argtypes = [_vp, C.POINTER(PsParticleArrs), _vp, _f, _f, _f, _f]
        L.ps_calc_forces.restype = C.c_int
        L.ps_prepare_particles.argtypes = [C.POINTER(PsParticleArrs), _vp, C.POINTER(_f),
                                           C.POINTER(_f), C.POINTER(_f)]
        L.ps_prepare_particles.restype = C.c_int
        L.ps_prepare_particles_ex.argtypes = [C.POINTER(PsParticleArrs), _vp, C.POINTER(_f),
                                              C.POINTER(_f), C.POINTER(_f), C.c_int]
        L.ps_prepare_particles_ex.restype = C.c_int
        L.ps_destroy.argtypes = [_vp]
        L.ps_destroy.restype = C.c_int
        L.ps_set_option.argtypes = [_vp, C.c_int, C.c_double]
        L.ps_set_option.restype = C.c_int
        L.ps_get_engine.argtypes = [_vp]
        L.ps_get_engine.restype = _vp
        L.ps_init_ex.argtypes = [C.POINTER(_vp), C.POINTER(PsConfigEx)]
        L.ps_init_ex.restype = C.c_int
        L.ps_device_scratch_bytes.argtypes = [_sz]
        L.ps_device_scratch_bytes.restype = _sz
        L.ps_calc_forces_async.argtypes = [_vp, C.POINTER(PsParticleArrs), _vp, _f, _f, _f, _f,
                                           _vp]
        L.ps_calc_forces_async.restype = C.c_int
        L.ps_calc_forces_wait.argtypes = [_vp]
        L.ps_calc_forces_wait.restype = C.c_int
        _host = L
    return _host


def check(where, rc):
    if rc != PS_OK:
        raise PsError(where, rc)
    return rc


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check("ps_hip_device_name", hip_lib().ps_hip_device_name(buf, 256))
    return buf.value.decode()


# ---------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------

class DeviceArray:
    """A 1-D device allocation of a numpy dtype (ps_hip_malloc)."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = hip_lib().ps_hip_malloc(max(self.nbytes, 1))
        if not p:
            raise PsError("ps_hip_malloc", PS_EOOM)
        self.ptr = p

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if self.nbytes:
            check("ps_hip_memcpy", hip_lib().ps_hip_memcpy(self.ptr, a.ctypes.data, self.nbytes, 1))

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        if self.nbytes:
            check("ps_hip_memcpy", hip_lib().ps_hip_memcpy(out.ctypes.data, self.ptr, self.nbytes, 2))
        return out

    def free(self):
        if self.ptr:
            hip_lib().ps_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return int(a)


# ---------------------------------------------------------------------------
# the reference API, mirrored
# ---------------------------------------------------------------------------

def make_arrs(x, y, z, mass, fx, fy, fz, ids, cnt) -> PsParticleArrs:
    return PsParticleArrs(_ptr(x), _ptr(y), _ptr(z), _ptr(mass), _ptr(fx), _ptr(fy), _ptr(fz),
                          _ptr(ids), int(cnt))


class Context:
    """ps_context_t plus the caller's buffer that hosts it (H:1705)."""

    def __init__(self, ptr, buff):
        self.ptr = ptr
        self._buff = buff

    @property
    def engine(self):
        return host_lib().ps_get_engine(self.ptr)


def ps_init(buff_size: int, max_particles: int, theta: float = 0.0, thrd_cnt: int = 4):
    """ps_init (H:1699-1767) with a host buffer of buff_size bytes.
    Returns (result, Context or None)."""
    buff = (C.c_uint8 * int(buff_size))()
    cfg = PsConfig(C.cast(buff, _vp), int(buff_size), int(max_particles), float(theta),
                   int(thrd_cnt))
    ctx = _vp()
    rc = host_lib().ps_init(C.byref(ctx), C.byref(cfg))
    return rc, (Context(ctx.value, buff) if rc == PS_OK else None)


def ps_init_ex(buff_size: int, max_particles: int, theta: float = 0.0, thrd_cnt: int = 4,
               dev_buff=None, dev_buff_size: int = 0):
    """ps_init_ex: ps_init with caller-owned device memory (dev_buff: a
    DeviceArray or device address).  Returns (result, Context or None)."""
    buff = (C.c_uint8 * int(buff_size))()
    base = PsConfig(C.cast(buff, _vp), int(buff_size), int(max_particles), float(theta),
                    int(thrd_cnt))
    cfg = PsConfigEx(base, _ptr(dev_buff), int(dev_buff_size))
    ctx = _vp()
    rc = host_lib().ps_init_ex(C.byref(ctx), C.byref(cfg))
    return rc, (Context(ctx.value, buff) if rc == PS_OK else None)


def device_scratch_bytes(max_particles: int) -> int:
    return int(host_lib().ps_device_scratch_bytes(int(max_particles)))


def alloc_stats():
    """(calls, bytes) of the library's own device and pinned allocations"""
    c, b = C.c_uint64(), C.c_uint64()
    check("ps_hip_alloc_stats", hip_lib().ps_hip_alloc_stats(C.byref(c), C.byref(b)))
    return c.value, b.value


FAIL_OFF, FAIL_ALLOC, FAIL_WAIT = 0, 1, 2


def debug_fail(what: int, after: int = 0):
    """Arm the library's failure hook: the allocation (FAIL_ALLOC) or host
    wait (FAIL_WAIT) after `after` others fails once; FAIL_OFF disarms."""
    check("ps_hip_debug_fail", hip_lib().ps_hip_debug_fail(int(what), int(after)))


class Stream:
    """A HIP stream from the library (ps_hip_stream_create)."""

    def __init__(self):
        h = _vp()
        check("ps_hip_stream_create", hip_lib().ps_hip_stream_create(C.byref(h)))
        self.h = h.value

    def synchronize(self):
        check("ps_hip_stream_synchronize", hip_lib().ps_hip_stream_synchronize(self.h))

    def close(self):
        if self.h:
            hip_lib().ps_hip_stream_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def ps_prepare_particles(arrs: PsParticleArrs, codes, mode: int | None = None):
    """ps_prepare_particles (H:1911-2170): returns (result, min, max, range)."""
    lo, hi, rg = _f(), _f(), _f()
    L = host_lib()
    if mode is None:
        rc = L.ps_prepare_particles(C.byref(arrs), _ptr(codes), C.byref(lo), C.byref(hi),
                                    C.byref(rg))
    else:
        rc = L.ps_prepare_particles_ex(C.byref(arrs), _ptr(codes), C.byref(lo), C.byref(hi),
                                       C.byref(rg), int(mode))
    return rc, lo.value, hi.value, rg.value


def ps_calc_forces(ctx: Context, arrs: PsParticleArrs, codes, cx, cy, cz, hw) -> int:
    """ps_calc_forces (H:1840-1909)."""
    return host_lib().ps_calc_forces(ctx.ptr if ctx else None, C.byref(arrs), _ptr(codes),
                                     float(cx), float(cy), float(cz), float(hw))


def ps_calc_forces_async(ctx: Context, arrs: PsParticleArrs, codes, cx, cy, cz, hw,
                         stream=None) -> int:
    """ps_calc_forces enqueued on a stream (a Stream, a handle, or None for
    the default stream) without waiting; complete with ps_calc_forces_wait."""
    h = stream.h if isinstance(stream, Stream) else stream
    return host_lib().ps_calc_forces_async(ctx.ptr if ctx else None, C.byref(arrs), _ptr(codes),
                                           float(cx), float(cy), float(cz), float(hw), h)


def ps_calc_forces_wait(ctx: Context) -> int:
    return host_lib().ps_calc_forces_wait(ctx.ptr if ctx else None)


def ps_destroy(ctx: Context) -> int:
    return host_lib().ps_destroy(ctx.ptr if ctx else None)


def ps_set_option(ctx: Context, key: int, value: float) -> int:
    return host_lib().ps_set_option(ctx.ptr, int(key), float(value))


def get_stats(ctx: Context) -> dict:
    st = PsStats()
    check("ps_hip_engine_get_stats", hip_lib().ps_hip_engine_get_stats(ctx.engine, C.byref(st)))
    return st.as_dict()


def export_tree(ctx: Context, cap: int) -> dict:
    """The last tick's octree in canonical level order (numpy)."""
    geom = np.zeros(4 * cap, np.float32)
    mp = np.zeros(4 * cap, np.float32)
    loc = np.zeros(4 * cap, np.float32)
    link = np.zeros(2 * cap, np.uint32)
    nend = np.zeros(cap, np.uint32)
    parent = np.zeros(cap, np.uint32)
    nn = C.c_uint32()
    check("ps_hip_engine_export_tree",
          hip_lib().ps_hip_engine_export_tree(ctx.engine, cap, geom.ctypes.data, mp.ctypes.data,
                                              loc.ctypes.data, link.ctypes.data, nend.ctypes.data,
                                              parent.ctypes.data, C.byref(nn)))
    n = nn.value
    link = link[: 2 * n].reshape(n, 2)
    leaf = (link[:, 1] & 0x80000000) != 0
    return dict(n_nodes=n, geom=geom[: 4 * n].reshape(n, 4), mp=mp[: 4 * n].reshape(n, 4),
                loc=loc[: 4 * n].reshape(n, 4), link=link, leaf=leaf, nend=nend[:n],
                parent=parent[:n],
                first=np.where(leaf, link[:, 0], 0).astype(np.uint32),
                mask=np.where(leaf, 0, link[:, 1] & 0xFF).astype(np.uint32))


def buff_size_for(n: int) -> int:
    """A host buffer that satisfies the reference's contract for n particles
    with room for 2.5 nodes per particle (80 B each) -- the same sizing the
    oracle's reference runs use."""
    return int(4 << 20) + 24 * int(n) + 200 * max(int(n), 1024)


class Solver:
    """Convenience wrapper: one context plus device-resident particle arrays,
    driven through the public API exactly as the reference's callers do
    (README.md:94-126)."""

    def __init__(self, n: int, theta: float = 2.0, buff_size: int | None = None,
                 traversal: int = TRAV_MT, kahan: bool = False, stats: bool = False,
                 device_scratch: bool = False):
        self.n = int(n)
        self.scratch = None
        if device_scratch:  # caller-owned device memory for every tick buffer
            self.scratch = DeviceArray(device_scratch_bytes(n), np.uint8)
            rc, self.ctx = ps_init_ex(buff_size or buff_size_for(n), n, theta,
                                      dev_buff=self.scratch, dev_buff_size=self.scratch.nbytes)
            check("ps_init_ex", rc)
        else:
            rc, self.ctx = ps_init(buff_size or buff_size_for(n), n, theta)
            check("ps_init", rc)
        if traversal != TRAV_MT:
            check("ps_set_option", ps_set_option(self.ctx, OPT_TRAVERSAL, traversal))
        if kahan:
            check("ps_set_option", ps_set_option(self.ctx, OPT_KAHAN, 1))
        if stats:
            check("ps_set_option", ps_set_option(self.ctx, OPT_STATS, 1))
        f32 = np.float32
        self.x, self.y, self.z, self.m = (DeviceArray(n, f32) for _ in range(4))
        self.fx, self.fy, self.fz = (DeviceArray(n, f32) for _ in range(3))
        self.id = DeviceArray(n, np.uint32)
        self.codes = DeviceArray(n, np.uint32)
        self.arrs = make_arrs(self.x, self.y, self.z, self.m, self.fx, self.fy, self.fz,
                              self.id, n)

    def load(self, x, y, z, m, ids=None):
        for d, a in ((self.x, x), (self.y, y), (self.z, z), (self.m, m)):
            d.upload(np.asarray(a, np.float32))
        self.id.upload(np.arange(self.n, dtype=np.uint32) if ids is None else ids)

    def reset_ids(self):
        check("ps_hip_step_reset_ids",
              hip_lib().ps_hip_step_reset_ids(self.ctx.engine, self.id.ptr, self.n))

    def frame_step(self, vel, tmp, dt):
        """the caller's frame step after a tick (README.md:94-161): gather the
        device velocity arrays vel = (vx, vy, vz) by id, semi-implicit Euler
        with time step dt, reset the ids; tmp: 3n device floats"""
        check("ps_hip_step_frame", hip_lib().ps_hip_step_frame(
            self.ctx.engine, self.id.ptr, vel[0].ptr, vel[1].ptr, vel[2].ptr, tmp.ptr,
            self.x.ptr, self.y.ptr, self.z.ptr, self.fx.ptr, self.fy.ptr, self.fz.ptr,
            self.m.ptr, dt, self.n))

    def tick(self, mode=None):
        """prepare + calc_forces with root = min + range/2 (README.md:104-126)."""
        rc, lo, hi, rg = ps_prepare_particles(self.arrs, self.codes, mode)
        check("ps_prepare_particles", rc)
        rc_ = np.float32(np.float32(lo) + np.float32(rg) / np.float32(2.0))
        hw = np.float32(rg) / np.float32(2.0)
        check("ps_calc_forces", ps_calc_forces(self.ctx, self.arrs, self.codes, rc_, rc_, rc_, hw))
        return lo, hi, rg

    def tick_async(self, stream=None, mode=None):
        """prepare (synchronous: the root comes from its bounding cube), then
        calc_forces enqueued on the stream; complete with wait()"""
        rc, lo, hi, rg = ps_prepare_particles(self.arrs, self.codes, mode)
        check("ps_prepare_particles", rc)
        rc_ = np.float32(np.float32(lo) + np.float32(rg) / np.float32(2.0))
        hw = np.float32(rg) / np.float32(2.0)
        check("ps_calc_forces_async", ps_calc_forces_async(self.ctx, self.arrs, self.codes, rc_,
                                                           rc_, rc_, hw, stream))
        return lo, hi, rg

    def wait(self):
        check("ps_calc_forces_wait", ps_calc_forces_wait(self.ctx))

    def result(self):
        return dict(x=self.x.numpy(), y=self.y.numpy(), z=self.z.numpy(), m=self.m.numpy(),
                    fx=self.fx.numpy(), fy=self.fy.numpy(), fz=self.fz.numpy(),
                    id=self.id.numpy(), codes=self.codes.numpy())

    def forces(self):
        return np.stack([self.fx.numpy(), self.fy.numpy(), self.fz.numpy()], 1)

    def stats(self):
        return get_stats(self.ctx)

    def tree(self):
        st = self.stats()
        return export_tree(self.ctx, max(int(st["n_nodes"]), 1))

    def direct_f64(self, positions):
        """FP64 direct sum on the device for the given sorted positions."""
        tg = np.ascontiguousarray(positions, np.uint32)
        out = np.zeros(3 * tg.size, np.float64)
        check("ps_hip_direct_f64",
              hip_lib().ps_hip_direct_f64(self.x.ptr, self.y.ptr, self.z.ptr, self.m.ptr, self.n,
                                          tg.ctypes.data, tg.size, out.ctypes.data))
        return out.reshape(-1, 3)

    def close(self):
        if self.ctx is not None:
            ps_destroy(self.ctx)
            self.ctx = None
        for a in (self.x, self.y, self.z, self.m, self.fx, self.fy, self.fz, self.id, self.codes):
            a.free()
        if self.scratch is not None:  # after ps_destroy: the engine no longer uses it
            self.scratch.free()
            self.scratch = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Many independent systems in one set of device arrays (ps_hip_batch_*):
    system k is particles [offsets[k], offsets[k+1]).  tick() sorts every
    segment in place and writes its forces, as separate ticks would."""

    def __init__(self, max_particles, max_systems, theta=2.0, traversal=TRAV_MT, graph=True,
                 kahan=False):
        L = hip_lib()
        h = C.c_void_p()
        check("ps_hip_batch_create", L.ps_hip_batch_create(C.byref(h), int(max_particles),
                                                           int(max_systems), float(theta),
                                                           int(traversal)))
        self.h = h
        self.max_particles = int(max_particles)
        if not graph:
            self.set_option(BATCH_OPT_GRAPH, 0)
        if kahan:
            self.set_option(OPT_KAHAN, 1)
        n = self.max_particles
        self.x, self.y, self.z, self.m = (DeviceArray(n, np.float32) for _ in range(4))
        self.fx, self.fy, self.fz = (DeviceArray(n, np.float32) for _ in range(3))
        self.id = DeviceArray(n, np.uint32)
        self.codes = DeviceArray(n, np.uint32)
        self.offsets = None

    def set_option(self, key, value):
        check("ps_hip_batch_set_option", hip_lib().ps_hip_batch_set_option(self.h, key, value))

    def set_leaf_caps(self, caps):
        """a leaf capacity per system (None: the OPT_LEAF_CAP option)"""
        if caps is None:
            check("ps_hip_batch_set_leaf_caps", hip_lib().ps_hip_batch_set_leaf_caps(self.h, None,
                                                                                     0))
            return
        c = np.ascontiguousarray(caps, np.uint8)
        check("ps_hip_batch_set_leaf_caps", hip_lib().ps_hip_batch_set_leaf_caps(
            self.h, c.ctypes.data, c.size))

    def load(self, systems):
        """systems: list of (x, y, z, m) float32 arrays; ids = index inside the system."""
        sizes = [len(t[0]) for t in systems]
        off = np.zeros(len(sizes) + 1, np.uint32)
        off[1:] = np.cumsum(sizes)
        n = int(off[-1])
        assert n <= self.max_particles
        cat = [np.concatenate([np.asarray(t[k], np.float32) for t in systems]) for k in range(4)]
        ids = np.concatenate([np.arange(s_, dtype=np.uint32) for s_ in sizes])
        self.n = n
        self.offsets = off
        self.host = dict(x=cat[0], y=cat[1], z=cat[2], m=cat[3], id=ids)
        self.restore()

    def restore(self):
        """upload the loaded (unsorted) state again"""
        h = self.host
        n = self.n
        for d, a in ((self.x, h["x"]), (self.y, h["y"]), (self.z, h["z"]), (self.m, h["m"]),
                     (self.id, h["id"])):
            check("ps_hip_memcpy", hip_lib().ps_hip_memcpy(d.ptr, a.ctypes.data, 4 * n, 1))

    def tick(self, mode=MORTON_TRUNC):
        o = np.ascontiguousarray(self.offsets, np.uint32)
        check("ps_hip_batch_tick", hip_lib().ps_hip_batch_tick(
            self.h, self.x.ptr, self.y.ptr, self.z.ptr, self.m.ptr, self.fx.ptr, self.fy.ptr,
            self.fz.ptr, self.id.ptr, self.codes.ptr, o.ctypes.data, len(o) - 1, int(mode)))

    def result(self):
        n = self.n
        out = {}
        for k, d in (("x", self.x), ("y", self.y), ("z", self.z), ("m", self.m),
                     ("fx", self.fx), ("fy", self.fy), ("fz", self.fz), ("id", self.id),
                     ("codes", self.codes)):
            out[k] = d.numpy()[:n]
        return out

    def system(self, res, k):
        a, b = int(self.offsets[k]), int(self.offsets[k + 1])
        return {key: v[a:b] for key, v in res.items()}

    def stats(self):
        st = PsBatchStats()
        check("ps_hip_batch_get_stats", hip_lib().ps_hip_batch_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in PsBatchStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            hip_lib().ps_hip_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
