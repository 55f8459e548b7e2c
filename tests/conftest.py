"""Shared test setup: paths, the `gpu` marker and cached workloads."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "turbogap_2.0_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

# the tests compare launch paths bit for bit: launch shapes must not depend on
# timing (tests/test_gpu_autotune.py turns the tuner on for its own contexts)
os.environ.pop("TGX_AUTOTUNE", None)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtgx.so)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.lib()
    return oracle
