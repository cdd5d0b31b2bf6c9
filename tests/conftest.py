import os
import sys
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def dev():
    """One libhgp context for the whole GPU session (tests run in one process)."""
    import axoncuda_amd as H
    d = H.Device(0)
    yield d
    d.close()


@pytest.fixture
def rng(request):
    """Seeded by the test's name (crc32: stable across processes, unlike hash())."""
    return np.random.default_rng(zlib.crc32(request.node.name.encode()))
