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
    relative error is ill-posed where a component crosses zero), taken over
    every component (all axes but the last; a 1-D array is one component).

    Strict about non-finite values: positions where a and b are equal
    (equal infinities included) count as no error; a NaN on either side, or
    an infinity the other side does not have, makes the error inf.  The
    denominator is the largest finite |b| of the component."""
    import numpy as np
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.shape != b.shape:
        raise ValueError(f"normwise: shapes differ, {a.shape} vs {b.shape}")
    if a.size == 0:
        return 0.0
    a2 = a.reshape(1, -1) if a.ndim <= 1 else a.reshape(-1, a.shape[-1])
    b2 = b.reshape(1, -1) if b.ndim <= 1 else b.reshape(-1, b.shape[-1])
    worst = 0.0
    for x, y in zip(a2, b2):
        differ = ~(x == y)
        if np.any(differ & ~(np.isfinite(x) & np.isfinite(y))):
            return float("inf")
        fy = np.abs(y[np.isfinite(y)])
        scale = max(float(fy.max()) if fy.size else 0.0, 1e-300)
        # only finite pairs differ here: equal infinities are masked out
        err = float(np.abs(x[differ] - y[differ]).max(initial=0.0)) / scale
        worst = max(worst, err)
    return worst


def cloud_sha(cloud):
    """SHA-256 over a cloud's arrays (the fixtures' cloud_sha)."""
    import hashlib
    import numpy as np
    h = hashlib.sha256()
    for a in (cloud.x, cloud.y, cloud.flag_1, cloud.flag_2, cloud.left, cloud.right,
              cloud.min_dist, cloud.conn_off, cloud.conn):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()
