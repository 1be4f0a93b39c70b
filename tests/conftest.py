import os
import sys

import pytest
# torch first: it brings its own HIP runtime, and libgsbn_amd.so then binds to
# that one (same soname).  Loaded the other way round, the process would hold
# two HIP runtimes, and device memory of one is not usable through the other.
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")
