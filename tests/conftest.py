import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def rel_l2(a, b):
    """Relative L2 difference per conservative variable (leading axis)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for m in range(a.shape[0]):
        den = np.linalg.norm(b[m])
        num = np.linalg.norm(a[m] - b[m])
        out.append(num / den if den > 0 else num)
    return np.array(out)


@pytest.fixture(scope="session")
def oracle():
    from oracle import ref
    ref.build()
    return ref
