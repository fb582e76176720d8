"""Shared test setup.  CPU tests (-m "not gpu") check the oracle against the
reference's golden vectors and known answers, and the C-ABI library's
exports; GPU tests (-m gpu) run the HIP path through the C-ABI and compare it
with the oracle."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels on an MI355X (gfx950)")


def golden_files(pattern=""):
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz") and pattern in f)


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O
