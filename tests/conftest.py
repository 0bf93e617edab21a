import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")
TINY_CASES = ["tiny_a", "tiny_b", "tiny_c", "tiny_d"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def have_gpu_device():
    return os.path.exists("/dev/kfd")


def load_tiny(name):
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz")))


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def plb():
    import porelb_amd

    porelb_amd.lib()
    return porelb_amd
