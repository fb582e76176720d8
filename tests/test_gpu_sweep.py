"""The reference benchmark's inputs (tests/benchmark.c: uniform [0,100)^3
from glibc rand() seed 7331, mass 1, PS_MULTITHREADING) through the drop-in
at its theta = 1.0 and at theta = 2.0, against the oracle: codes, ids,
tree and multipoles bit for bit, M2L and P2P pair counts exact, forces
within the parity tolerance.  At theta = 1.0 the depth-2 traversal roots
let a target take the M2L of a source that contains it (the corner depth-2
cells against the root); the GPU must reproduce those lists too."""
import pytest

from gpu_helpers import assert_tick_parity, gpu_tick, log
from polesitter_amd import api
from polesitter_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("theta", [1.0, 2.0])
@pytest.mark.parametrize("n", [10000, 50000])
def test_benchmark_inputs_match_oracle(oracle, n, theta):
    x, y, z, m = W.benchmark_cube(n)
    o = oracle.tick(x, y, z, m, theta=theta, mode=api.MORTON_AVX2)
    g = gpu_tick(x, y, z, m, theta=theta, mode=api.MORTON_AVX2)
    f, l = assert_tick_parity(g, o)
    log(f"benchmark cube n={n} theta={theta}: m2l {g['stats']['m2l_pairs']} "
        f"p2p {g['stats']['p2p_pairs']} force rms {f:.2e}")
