import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hypergrey-sih-26_amd")
ORACLE = os.path.join(ROOT, "oracle")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libhgqp.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import Oracle
    return Oracle("oracle")


@pytest.fixture(scope="session")
def reference():
    from oracle import Oracle, available
    if not available("reference"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return Oracle("reference")


@pytest.fixture(scope="session")
def hg():
    import hgqp
    hgqp.lib()
    return hgqp
