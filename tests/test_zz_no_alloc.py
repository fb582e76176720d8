"""ps_hip_prepare_noalloc: ps_hip_prepare with its scratch in a static device
array instead of library allocations.  Its results must equal the existing
entry point's bit for bit (codes, min, max, range) on consecutive calls, the
second one smaller, in both quantisation modes, and it must refuse host
arrays and sizes beyond its stated maximum."""
import ctypes as C

import numpy as np
import pytest

from gpu_helpers import log
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu


def prepare(fn, x, y, z, codes, n, mode):
    lo, hi, rg = C.c_float(), C.c_float(), C.c_float()
    rc = fn(x.ptr, y.ptr, z.ptr, n, codes.ptr, C.byref(lo), C.byref(hi), C.byref(rg), mode)
    return rc, (lo.value, hi.value, rg.value), codes.numpy()[:n]


@pytest.mark.parametrize("mode", [api.MORTON_TRUNC, api.MORTON_AVX2])
def test_noalloc_prepare_equals_prepare(mode):
    L = api.hip_lib()
    sizes = (300_001, 4_099)  # the second call smaller: the scratch keeps the first's data
    cap = max(sizes)
    dx, dy, dz = (api.DeviceArray(cap, np.float32) for _ in range(3))
    c_ref = api.DeviceArray(cap, np.uint32)
    c_new = api.DeviceArray(cap, np.uint32)
    for k, n in enumerate(sizes):
        x, y, z, _ = W.plummer(n, seed=200 + k)
        for d, a in ((dx, x), (dy, y), (dz, z)):
            d.upload(np.concatenate([a, np.zeros(cap - n, np.float32)]))
        log(f"noalloc prepare: call {k + 1}, n={n}, mode={mode}")
        rc0, p0, codes0 = prepare(L.ps_hip_prepare, dx, dy, dz, c_ref, n, mode)
        rc1, p1, codes1 = prepare(L.ps_hip_prepare_noalloc, dx, dy, dz, c_new, n, mode)
        assert rc0 == api.PS_OK and rc1 == api.PS_OK
        assert np.array_equal(np.float32(p0), np.float32(p1)), (p0, p1)
        assert np.array_equal(codes0, codes1)
        assert codes1.max() < (1 << 30)


def test_noalloc_prepare_rejects_what_it_cannot_do():
    L = api.hip_lib()
    n = 1000
    x, y, z, _ = W.uniform_cube(n, seed=5)
    dx, dy, dz = (api.DeviceArray.from_numpy(a) for a in (x, y, z))
    codes = api.DeviceArray(n, np.uint32)
    hx = np.ascontiguousarray(x)
    hc = np.zeros(n, np.uint32)
    lo, hi, rg = C.c_float(), C.c_float(), C.c_float()
    args = (C.byref(lo), C.byref(hi), C.byref(rg), api.MORTON_TRUNC)
    # host arrays would need staging buffers, i.e. an allocation
    assert L.ps_hip_prepare_noalloc(hx.ctypes.data, dy.ptr, dz.ptr, n, codes.ptr,
                                    *args) == api.PS_EINVAL
    assert L.ps_hip_prepare_noalloc(dx.ptr, dy.ptr, dz.ptr, n, hc.ctypes.data,
                                    *args) == api.PS_EINVAL
    # beyond the stated maximum, and the reference's own argument errors
    assert L.ps_hip_prepare_noalloc(dx.ptr, dy.ptr, dz.ptr, 1 << 32, codes.ptr,
                                    *args) == api.PS_EINVAL
    assert L.ps_hip_prepare_noalloc(dx.ptr, dy.ptr, dz.ptr, 0, codes.ptr,
                                    *args) == api.PS_EINVAL
    assert L.ps_hip_prepare_noalloc(dx.ptr, dy.ptr, dz.ptr, n, codes.ptr, C.byref(lo),
                                    C.byref(hi), C.byref(rg), 7) == api.PS_EINVAL
    assert L.ps_hip_prepare_noalloc(dx.ptr, dy.ptr, dz.ptr, n, codes.ptr, *args) == api.PS_OK
