import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    if not oracle.available("restatement"):
        oracle.build(reference=False)
    return oracle


@pytest.fixture(scope="session")
def reference_available(oracle_mod):
    return oracle_mod.available("reference")
