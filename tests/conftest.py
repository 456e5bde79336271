import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    # the autotuner's JSON cache of this test session, apart from the user's
    if "QGW_TUNE_CACHE" not in os.environ:
        import tempfile
        os.environ["QGW_TUNE_CACHE"] = os.path.join(tempfile.mkdtemp(prefix="qgw_tune_"), "tune.json")
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libqgw_hip.so)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def rel(a, b):
    import numpy as np
    d = np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel())
    s = np.linalg.norm(np.asarray(b).ravel())
    return d / s if s > 0 else d
