import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# The oracle's products are small (n <= 1024): one BLAS thread runs them several times faster
# than a threaded BLAS that spins up all host cores for each call.
try:
    from threadpoolctl import threadpool_limits
    _BLAS_LIMIT = threadpool_limits(1)
except ImportError:          # threadpoolctl absent: the BLAS keeps its own default
    _BLAS_LIMIT = None


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libalfq.so")
    config.addinivalue_line("markers", "slow: takes more than a few seconds")
