import os
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "yocto-mirrors_gitlab_kleidi_kleidicv_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: full-size (4K and larger) cases")


@pytest.fixture
def rng(request):
    """Deterministic generator seeded from the test's node id."""
    return np.random.default_rng(zlib.crc32(request.node.nodeid.encode()))
