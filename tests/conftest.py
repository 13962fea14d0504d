import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "c-_serial_amd"), REPO):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: takes more than ~30 s")


def golden(name):
    import numpy as np
    path = os.path.join(GOLDEN, name)
    if not os.path.isfile(path):
        pytest.skip(f"fixture {name} not generated")
    return np.load(path, allow_pickle=False)


def normwise(a, b):
    """max|a-b| / max|b| per leading component (the parity measure: pointwise
    relative error is ill-posed where a component crosses zero)."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 1:
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    a2 = a.reshape(-1, a.shape[-1])
    b2 = b.reshape(-1, b.shape[-1])
    return max(float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)) for x, y in zip(a2, b2))
