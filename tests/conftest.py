import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gd():
    import gdiamond_amd
    gdiamond_amd.lib()
    return gdiamond_amd


@pytest.fixture(scope="session")
def gpu(gd):
    if gd.device_count() < 1:
        pytest.fail("no HIP device visible for a test marked gpu")
    return gd


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_cases():
    return sorted(f for f in os.listdir(GOLDEN) if f.startswith("ref_seq_") and f.endswith(".npz"))
