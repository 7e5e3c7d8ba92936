import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "bem-cpp_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: oracle cases that take tens of seconds on the CPU")


def golden(name):
    return np.load(os.path.join(GOLDEN, name))


def python_to_sorted(keys):
    """Index array mapping the sorted (std::map, rwg.cpp:8-40) RWG order onto the
    Python reference's insertion order (bem_core.py:399-412): entry i of the
    sorted order is Python RWG ``order[i]``."""
    keys = np.asarray(keys)
    return np.lexsort((keys[:, 1], keys[:, 0]))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def bem():
    import bem_amd
    return bem_amd
