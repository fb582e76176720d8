"""The reference benchmark's inputs (tests/benchmark.c:158-165): glibc
rand() after srand(7331), scaled by ((float)rand() / RAND_MAX) * 100.0F.
Pinned by the C library's own rand() and by the benchmark's expression
compiled with gcc."""
import ctypes
import subprocess

import numpy as np

from polesitter_amd import workloads as W


def test_glibc_rand_matches_the_c_library():
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    for seed in (7331, 1, 0, 123456789):
        libc.srand(ctypes.c_uint(seed))
        want = np.array([libc.rand() for _ in range(5000)], np.uint32)
        assert np.array_equal(W.glibc_rand(seed, 5000), want), seed


def test_benchmark_cube_matches_the_compiled_expression(tmp_path):
    src = tmp_path / "b.c"
    src.write_text('#include <stdio.h>\n#include <stdlib.h>\n'
                   'int main(void){ srand(7331); for (int i = 0; i < 3000; i++) {\n'
                   ' float x = ((float)rand() / RAND_MAX) * 100.0F;\n'
                   ' fwrite(&x, 4, 1, stdout); } return 0; }\n')
    exe = tmp_path / "b"
    subprocess.run(["gcc", "-std=c99", "-O2", str(src), "-o", str(exe)], check=True)
    raw = subprocess.run([str(exe)], capture_output=True, check=True).stdout
    want = np.frombuffer(raw, np.float32).reshape(1000, 3)
    x, y, z, m = W.benchmark_cube(1000)
    assert np.array_equal(np.stack([x, y, z], 1).view(np.uint32), want.view(np.uint32))
    assert np.all(m == 1.0) and x.min() >= 0.0 and x.max() <= 100.0
