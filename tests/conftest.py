"""Test configuration: markers, import paths and the oracle build.

Tests marked ``gpu`` need an MI355X and the in-tree HIP library; everything else
runs on the CPU (oracle against the reference's golden values, the compiled
reference checker, ABI exports, scene construction).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "adaptive-pf-flip_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X device and the in-tree HIP library")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    import pyoracle
    if not os.path.exists(pyoracle.ORACLE_SO):
        pyoracle.build(ref=False)
    yield
