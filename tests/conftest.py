import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    # the CPU oracle is the checker; build it if this checkout has not yet
    lib = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


def rel_err(a, b, order):
    """Per-variable normwise relative error ||a-b|| / ||b|| (columns)."""
    import numpy as np
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    if order == "inf":
        num, den = d.max(0), np.abs(b).max(0)
    else:
        num, den = d.sum(0), np.abs(b).sum(0)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
