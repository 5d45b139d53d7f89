import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: longer CPU runs")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "reference_vectors.json")) as f:
        return json.load(f)


def _ensure_oracle_built():
    import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    return pyoracle


@pytest.fixture(scope="session")
def po():
    return _ensure_oracle_built()


@pytest.fixture(scope="session")
def oracle70(po):
    return po.Oracle(degree=70)


@pytest.fixture()
def oracle(po):
    """A fresh oracle handle without datasets (analytic forces only)."""
    return po.Oracle(degree=None)
