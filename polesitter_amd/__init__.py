"""polesitter_amd -- the polesitter FMM tick on AMD MI355X (gfx950).

The product is the C library pair built by ``polesitter_amd.build``:
libpolesitter_hip.so (HIP kernels + the ps_hip_* C-ABI) and libpolesitter.so
(the C host layer of include/polesitter.h).  ``polesitter_amd.api`` mirrors
the reference's C API in Python over ctypes; ``polesitter_amd.workloads``
generates the synthetic particle sets.
"""
import os as _os

LIBDIR = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "lib")


def built() -> bool:
    return all(_os.path.exists(_os.path.join(LIBDIR, f))
               for f in ("libpolesitter_hip.so", "libpolesitter.so"))
