"""Shared test setup: repository paths, the `gpu` marker and the oracle build."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and hussar_amd/libhussar_hip.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle_bindings as ob
    ob.lib()
    return ob


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
